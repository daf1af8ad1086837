// URDF reader and model compiler (host side). See model.h.
#include "model.h"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <memory>
#include <sstream>

namespace sapamd {

// ------------------------------------------------------------------ transforms
Xform xmul(const Xform &a, const Xform &b) {
  Xform o;
  for (int i = 0; i < 3; ++i) {
    for (int j = 0; j < 3; ++j)
      o.R[3 * i + j] = a.R[3 * i] * b.R[j] + a.R[3 * i + 1] * b.R[3 + j] + a.R[3 * i + 2] * b.R[6 + j];
    o.p[i] = a.R[3 * i] * b.p[0] + a.R[3 * i + 1] * b.p[1] + a.R[3 * i + 2] * b.p[2] + a.p[i];
  }
  return o;
}

static void rpy_matrix(double r, double p, double y, double R[9]) {
  // URDF fixed-axis roll, pitch, yaw: R = Rz(y) Ry(p) Rx(r)
  double cr = std::cos(r), sr = std::sin(r), cp = std::cos(p), sp = std::sin(p), cy = std::cos(y),
         sy = std::sin(y);
  R[0] = cy * cp; R[1] = cy * sp * sr - sy * cr; R[2] = cy * sp * cr + sy * sr;
  R[3] = sy * cp; R[4] = sy * sp * sr + cy * cr; R[5] = sy * sp * cr - cy * sr;
  R[6] = -sp;     R[7] = cp * sr;                R[8] = cp * cr;
}

// ------------------------------------------------------------------ inertia
static void outer_shift(double m, const double c[3], double out[9], double sign) {
  double cc = c[0] * c[0] + c[1] * c[1] + c[2] * c[2];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) out[3 * i + j] += sign * m * ((i == j ? cc : 0.0) - c[i] * c[j]);
}

BodyInertia BodyInertia::from_com(double m, const double c[3], const double Ic[9]) {
  BodyInertia b;
  b.m = m;
  for (int i = 0; i < 3; ++i) b.h[i] = m * c[i];
  std::memcpy(b.IO, Ic, sizeof(b.IO));
  outer_shift(m, c, b.IO, 1.0);
  return b;
}

void BodyInertia::com_form(double &mass, double c[3], double Ic[9]) const {
  mass = m;
  for (int i = 0; i < 3; ++i) c[i] = m > 0 ? h[i] / m : 0.0;
  std::memcpy(Ic, IO, sizeof(IO));
  outer_shift(m, c, Ic, -1.0);
}

BodyInertia BodyInertia::moved(const Xform &X) const {
  double mass, c[3], Ic[9];
  com_form(mass, c, Ic);
  double c2[3], RI[9], I2[9];
  for (int i = 0; i < 3; ++i)
    c2[i] = X.R[3 * i] * c[0] + X.R[3 * i + 1] * c[1] + X.R[3 * i + 2] * c[2] + X.p[i];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      RI[3 * i + j] = X.R[3 * i] * Ic[j] + X.R[3 * i + 1] * Ic[3 + j] + X.R[3 * i + 2] * Ic[6 + j];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      I2[3 * i + j] = RI[3 * i] * X.R[3 * j] + RI[3 * i + 1] * X.R[3 * j + 1] + RI[3 * i + 2] * X.R[3 * j + 2];
  return from_com(mass, c2, I2);
}

void BodyInertia::add(const BodyInertia &o) {
  m += o.m;
  for (int i = 0; i < 3; ++i) h[i] += o.h[i];
  for (int i = 0; i < 9; ++i) IO[i] += o.IO[i];
}

// ------------------------------------------------------------------ minimal XML DOM
namespace {
struct XNode {
  std::string tag;
  std::map<std::string, std::string> attr;
  std::vector<std::unique_ptr<XNode>> kids;
  const char *get(const char *k) const {
    auto it = attr.find(k);
    return it == attr.end() ? nullptr : it->second.c_str();
  }
  const XNode *child(const char *t) const {
    for (auto &k : kids)
      if (k->tag == t) return k.get();
    return nullptr;
  }
};

class XmlParser {
 public:
  explicit XmlParser(const std::string &s) : s_(s) {}
  std::unique_ptr<XNode> parse() {
    skip_misc();
    auto root = element();
    return root;
  }

 private:
  const std::string &s_;
  size_t i_ = 0;
  [[noreturn]] void fail(const char *what) {
    std::ostringstream os;
    os << "URDF parse error at offset " << i_ << ": " << what;
    throw std::invalid_argument(os.str());
  }
  bool starts(const char *t) const { return s_.compare(i_, std::strlen(t), t) == 0; }
  void ws() {
    while (i_ < s_.size() && std::isspace((unsigned char)s_[i_])) ++i_;
  }
  void skip_until(const char *t) {
    size_t k = s_.find(t, i_);
    if (k == std::string::npos) fail("unterminated construct");
    i_ = k + std::strlen(t);
  }
  void skip_misc() {
    for (;;) {
      ws();
      if (starts("<?")) skip_until("?>");
      else if (starts("<!--")) skip_until("-->");
      else if (starts("<!DOCTYPE")) skip_until(">");
      else return;
    }
  }
  std::string name() {
    size_t b = i_;
    while (i_ < s_.size() && (std::isalnum((unsigned char)s_[i_]) || std::strchr("_-.:", s_[i_]))) ++i_;
    if (b == i_) fail("expected a name");
    return s_.substr(b, i_ - b);
  }
  static std::string unescape(const std::string &v) {
    std::string o;
    for (size_t k = 0; k < v.size(); ++k) {
      if (v[k] != '&') { o += v[k]; continue; }
      size_t e = v.find(';', k);
      std::string ent = v.substr(k + 1, e - k - 1);
      if (ent == "amp") o += '&';
      else if (ent == "lt") o += '<';
      else if (ent == "gt") o += '>';
      else if (ent == "quot") o += '"';
      else if (ent == "apos") o += '\'';
      else if (!ent.empty() && ent[0] == '#') o += (char)std::strtol(ent.c_str() + (ent[1] == 'x' ? 2 : 1), nullptr, ent[1] == 'x' ? 16 : 10);
      k = e;
    }
    return o;
  }
  std::unique_ptr<XNode> element() {
    if (i_ >= s_.size() || s_[i_] != '<') fail("expected an element");
    ++i_;
    auto n = std::make_unique<XNode>();
    n->tag = name();
    for (;;) {
      ws();
      if (starts("/>")) { i_ += 2; return n; }
      if (starts(">")) { ++i_; break; }
      std::string k = name();
      ws();
      if (i_ >= s_.size() || s_[i_] != '=') fail("expected '='");
      ++i_;
      ws();
      char qc = s_[i_];
      if (qc != '"' && qc != '\'') fail("expected a quoted attribute value");
      size_t e = s_.find(qc, i_ + 1);
      if (e == std::string::npos) fail("unterminated attribute");
      n->attr[k] = unescape(s_.substr(i_ + 1, e - i_ - 1));
      i_ = e + 1;
    }
    for (;;) {
      size_t lt = s_.find('<', i_);
      if (lt == std::string::npos) fail("unterminated element");
      i_ = lt;
      if (starts("<!--")) { skip_until("-->"); continue; }
      if (starts("<![CDATA[")) { skip_until("]]>"); continue; }
      if (starts("<?")) { skip_until("?>"); continue; }
      if (starts("</")) {
        i_ += 2;
        std::string t = name();
        if (t != n->tag) fail("mismatched closing tag");
        ws();
        if (i_ >= s_.size() || s_[i_] != '>') fail("expected '>'");
        ++i_;
        return n;
      }
      n->kids.push_back(element());
    }
  }
};

std::vector<double> numbers(const char *s, size_t n, std::initializer_list<double> dflt) {
  std::vector<double> v;
  if (!s) return std::vector<double>(dflt);
  std::istringstream is(s);
  double x;
  while (is >> x) v.push_back(x);
  if (v.size() != n) throw std::invalid_argument(std::string("URDF: expected ") + std::to_string(n) + " numbers in '" + s + "'");
  return v;
}

Xform origin_of(const XNode *o) {
  Xform X;
  if (!o) return X;
  auto xyz = numbers(o->get("xyz"), 3, {0, 0, 0});
  auto rpy = numbers(o->get("rpy"), 3, {0, 0, 0});
  rpy_matrix(rpy[0], rpy[1], rpy[2], X.R);
  for (int i = 0; i < 3; ++i) X.p[i] = xyz[i];
  return X;
}

double attr_num(const XNode *n, const char *k, double dflt) {
  const char *v = n ? n->get(k) : nullptr;
  return v ? std::strtod(v, nullptr) : dflt;
}

struct UJ {
  std::string name, type, parent, child;
  Xform origin;
  double axis[3] = {1, 0, 0};
  double lower = 0, upper = 0;
};
}  // namespace

// ------------------------------------------------------------------ model
ArticulatedModel::ArticulatedModel(const std::string &urdf, const double gravity[3]) {
  for (int i = 0; i < 3; ++i) gravity_[i] = gravity[i];
  auto root = XmlParser(urdf).parse();
  if (root->tag != "robot") throw std::invalid_argument("URDF root element must be <robot>");

  std::map<std::string, BodyInertia> link_inertia;
  std::map<std::string, std::array<double, 3>> link_com;
  std::vector<std::string> link_names;
  for (auto &k : root->kids) {
    if (k->tag != "link") continue;
    const char *nm = k->get("name");
    if (!nm) throw std::invalid_argument("URDF link without a name");
    BodyInertia bi;
    if (const XNode *in = k->child("inertial")) {
      Xform T = origin_of(in->child("origin"));
      double m = attr_num(in->child("mass"), "value", 0.0);
      const XNode *ii = in->child("inertia");
      double I[9] = {attr_num(ii, "ixx", 0), attr_num(ii, "ixy", 0), attr_num(ii, "ixz", 0),
                     attr_num(ii, "ixy", 0), attr_num(ii, "iyy", 0), attr_num(ii, "iyz", 0),
                     attr_num(ii, "ixz", 0), attr_num(ii, "iyz", 0), attr_num(ii, "izz", 0)};
      // inertial origin rotation rotates the tensor: I = R I_urdf R^T (about the COM)
      double RI[9], Ir[9];
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
          RI[3 * i + j] = T.R[3 * i] * I[j] + T.R[3 * i + 1] * I[3 + j] + T.R[3 * i + 2] * I[6 + j];
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
          Ir[3 * i + j] = RI[3 * i] * T.R[3 * j] + RI[3 * i + 1] * T.R[3 * j + 1] + RI[3 * i + 2] * T.R[3 * j + 2];
      bi = BodyInertia::from_com(m, T.p, Ir);
      link_com[nm] = {T.p[0], T.p[1], T.p[2]};
    } else {
      link_com[nm] = {0.0, 0.0, 0.0};
    }
    link_inertia[nm] = bi;
    link_names.push_back(nm);
  }
  // joints, keyed by name: urdfdom keeps them in a std::map and links its
  // children in that (sorted) order, which fixes the depth-first joint order.
  std::map<std::string, UJ> ujoints;
  for (auto &k : root->kids) {
    if (k->tag != "joint") continue;
    UJ j;
    const char *nm = k->get("name");
    const char *tp = k->get("type");
    const XNode *pa = k->child("parent"), *ch = k->child("child");
    if (!nm || !tp || !pa || !ch || !pa->get("link") || !ch->get("link"))
      throw std::invalid_argument("URDF joint missing name/type/parent/child");
    j.name = nm;
    j.type = tp;
    j.parent = pa->get("link");
    j.child = ch->get("link");
    j.origin = origin_of(k->child("origin"));
    if (const XNode *ax = k->child("axis")) {
      auto a = numbers(ax->get("xyz"), 3, {1, 0, 0});
      for (int i = 0; i < 3; ++i) j.axis[i] = a[i];
    }
    if (const XNode *lim = k->child("limit")) {
      j.lower = attr_num(lim, "lower", 0);
      j.upper = attr_num(lim, "upper", 0);
    }
    if (!link_inertia.count(j.parent) || !link_inertia.count(j.child))
      throw std::invalid_argument("URDF joint " + j.name + " references an unknown link");
    ujoints[j.name] = j;
  }
  std::map<std::string, std::vector<const UJ *>> children;
  std::map<std::string, int> is_child;
  for (auto &kv : ujoints) {
    children[kv.second.parent].push_back(&kv.second);
    if (is_child[kv.second.child]++) throw std::invalid_argument("URDF link with two parent joints: " + kv.second.child);
  }
  std::string root_link;
  int nroots = 0;
  for (auto &ln : link_names)
    if (!is_child.count(ln)) { root_link = ln; ++nroots; }
  if (nroots != 1) throw std::invalid_argument("URDF must have exactly one root link");

  const auto with_com = [&](Frame f) {
    for (int i = 0; i < 3; ++i) f.com[i] = link_com[f.name][i];
    return f;
  };
  frames_.push_back(with_com(Frame{root_link, -1, Xform{}}));
  frame_index_[root_link] = 0;
  std::function<void(const std::string &)> visit = [&](const std::string &link) {
    for (const UJ *uj : children[link]) {
      const Frame pf = frames_[frame_index_[link]];
      Xform place = xmul(pf.place, uj->origin);
      if (uj->type == "fixed") {
        frame_index_[uj->child] = (int)frames_.size();
        frames_.push_back(with_com(Frame{uj->child, pf.joint, place}));
        if (pf.joint >= 0) joints_[pf.joint].inertia.add(link_inertia[uj->child].moved(place));
      } else {
        Joint j;
        j.name = uj->name;
        j.parent = pf.joint;
        j.place = place;
        j.inertia = link_inertia[uj->child];
        j.lower = uj->lower;
        j.upper = uj->upper;
        j.depth = pf.joint >= 0 ? joints_[pf.joint].depth + 1 : 1;
        double n = std::sqrt(uj->axis[0] * uj->axis[0] + uj->axis[1] * uj->axis[1] + uj->axis[2] * uj->axis[2]);
        if (uj->type == "revolute") j.kind = JK_REVOLUTE;
        else if (uj->type == "prismatic") j.kind = JK_PRISMATIC;
        else if (uj->type == "continuous") { j.kind = JK_CONTINUOUS; j.nq = 2; }
        else if (uj->type == "floating") { j.kind = JK_UNSUPPORTED; j.nq = 7; j.nv = 6; }
        else if (uj->type == "planar") { j.kind = JK_UNSUPPORTED; j.nq = 4; j.nv = 3; }
        else throw std::invalid_argument("URDF joint " + uj->name + " has unknown type " + uj->type);
        if (j.kind != JK_UNSUPPORTED) {
          if (!(n > 0)) throw std::invalid_argument("URDF joint " + uj->name + " has a zero axis");
          for (int i = 0; i < 3; ++i) j.axis[i] = uj->axis[i] / n;
        } else {
          unsupported_.push_back(uj->name);
        }
        int id = (int)joints_.size();
        joints_.push_back(j);
        nv_ += j.nv;
        nq_ += j.nq;
        frame_index_[uj->child] = (int)frames_.size();
        frames_.push_back(with_com(Frame{uj->child, id, Xform{}}));
      }
      visit(uj->child);
    }
  };
  visit(root_link);
  vuser_.assign(joints_.size(), -1);
}

void ArticulatedModel::set_joint_order(const std::vector<std::string> &names) {
  std::vector<int> ids;
  int count = 0;
  for (auto &nm : names) {
    int id = -2;
    if (nm == "universe") id = -1;
    for (size_t j = 0; j < joints_.size(); ++j)
      if (joints_[j].name == nm) id = (int)j;
    if (id == -2) throw std::invalid_argument("invalid names in setJointOrder");
    ids.push_back(id);
    if (id >= 0) count += joints_[id].nv;
  }
  if (count != nv_) throw std::runtime_error("setJointOrder failed");
  std::vector<int> seen(joints_.size(), 0);
  for (int id : ids)
    if (id >= 0 && seen[id]++) throw std::runtime_error("setJointOrder failed");
  std::vector<int> vu(joints_.size(), -1);
  int k = 0;
  for (int id : ids) {
    if (id < 0) continue;
    vu[id] = k;
    k += joints_[id].nv;
  }
  vuser_ = vu;
  user_joint_seq_ = ids;
  order_set_ = true;
}

void ArticulatedModel::set_link_order(const std::vector<std::string> &names) {
  std::vector<int> v;
  for (auto &nm : names) {
    auto it = frame_index_.find(nm);
    if (it == frame_index_.end()) throw std::invalid_argument("invalid names in setLinkOrder");
    v.push_back(it->second);
  }
  link_frames_ = v;
}

void ArticulatedModel::check_compute(size_t qlen) const {
  // posS2P (pinocchio_model.cpp:49-72): walks the user joints, throwing on
  // nq > 2 joints, then requires the user vector to cover exactly the dofs.
  size_t count = 0;
  for (int id : user_joint_seq_) {
    if (id < 0) continue;
    if (joints_[id].kind == JK_UNSUPPORTED)
      throw std::runtime_error("Unsupported joint in computation. Currently support: fixed, revolute, prismatic");
    count += joints_[id].nv;
  }
  if (count != qlen) throw std::runtime_error("posS2P failed");
}

std::vector<int> ArticulatedModel::support(int joint) const {
  std::vector<int> s;
  for (int j = joint; j >= 0; j = joints_[j].parent) s.push_back(j);
  std::reverse(s.begin(), s.end());
  return s;
}

int ArticulatedModel::max_depth() const {
  int d = 0;
  for (auto &j : joints_) d = std::max(d, j.depth);
  return d;
}

bool ArticulatedModel::is_chain() const {
  for (size_t j = 0; j < joints_.size(); ++j)
    if (joints_[j].parent != (int)j - 1) return false;
  return true;
}

}  // namespace sapamd
