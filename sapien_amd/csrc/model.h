// Host-side articulated-body model: URDF text -> joint/frame tables with the
// semantics of the reference's model construction (pinocchio_model.cpp:39-47,
// which calls pinocchio::urdf::buildModelFromXML), plus the user-order state of
// set_joint_order / set_link_order (pinocchio_model.cpp:99-140).
#pragma once
#include <array>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

namespace sapamd {

// Joint kinds supported by computation (posS2P, pinocchio_model.cpp:53-67):
// revolute and prismatic (nq = nv = 1), continuous (nq = 2 as cos/sin, nv = 1).
// floating/planar joints parse but make every compute call fail with the
// reference's error text.
enum JointKind : int { JK_REVOLUTE = 0, JK_PRISMATIC = 1, JK_CONTINUOUS = 2, JK_UNSUPPORTED = 3 };

struct Xform {  // rigid transform: x_parent = R x_child + p, R row-major
  double R[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
  double p[3] = {0, 0, 0};
};
Xform xmul(const Xform &a, const Xform &b);

// Spatial inertia in additive form about the frame origin: mass, first moment
// h = m c, rotational inertia I_O = I_c + m (|c|^2 I - c c^T).
struct BodyInertia {
  double m = 0, h[3] = {0, 0, 0}, IO[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  static BodyInertia from_com(double m, const double c[3], const double Ic[9]);
  BodyInertia moved(const Xform &X) const;  // expressed in the parent frame of X
  void add(const BodyInertia &o);
  void com_form(double &m, double c[3], double Ic[9]) const;
};

struct Joint {
  std::string name;
  int parent = -1;  // parent joint (-1 = universe)
  int kind = JK_REVOLUTE;
  int nq = 1, nv = 1;
  double axis[3] = {1, 0, 0};
  Xform place;  // parent-joint frame -> this joint frame (fixed joints folded in)
  BodyInertia inertia;
  double lower = 0, upper = 0;
  int depth = 0;
};

struct Frame {
  std::string name;
  int joint = -1;  // joint the frame is attached to (-1 = universe)
  Xform place;     // joint frame -> link frame
  double com[3] = {0, 0, 0};  // the link's own centre of mass in the link frame (its inertial origin)
};

class ArticulatedModel {
 public:
  ArticulatedModel(const std::string &urdf, const double gravity[3]);

  // Reference semantics: unknown names -> std::invalid_argument, dof count
  // mismatch -> std::runtime_error("setJointOrder failed").
  void set_joint_order(const std::vector<std::string> &names);
  void set_link_order(const std::vector<std::string> &names);

  int nv() const { return nv_; }
  int nq() const { return nq_; }
  int njoints() const { return (int)joints_.size(); }
  int nlinks() const { return (int)link_frames_.size(); }
  bool joint_order_set() const { return order_set_; }
  bool supported() const { return unsupported_.empty(); }
  const std::vector<Joint> &joints() const { return joints_; }
  const std::vector<Frame> &frames() const { return frames_; }
  const std::vector<int> &link_frames() const { return link_frames_; }
  const std::vector<int> &vuser() const { return vuser_; }  // per joint: user dof index
  const double *gravity() const { return gravity_; }
  // the reference's posS2P pre-checks, thrown with its messages
  void check_compute(size_t qlen) const;
  // joints from the root down to `joint` (inclusive), internal indices
  std::vector<int> support(int joint) const;
  int max_depth() const;
  bool is_chain() const;  // every joint's parent is the previous joint

 private:
  std::vector<Joint> joints_;
  std::vector<Frame> frames_;
  std::map<std::string, int> frame_index_;
  std::vector<int> link_frames_;
  std::vector<int> vuser_;
  std::vector<std::string> unsupported_;
  std::vector<int> user_joint_seq_;  // joint ids in set_joint_order order
  int nv_ = 0, nq_ = 0;
  bool order_set_ = false;
  double gravity_[3];
};

}  // namespace sapamd
