// ORACLE — test infrastructure only. Not part of the product: only tests/, the
// smoke check and the CPU-baseline leg of bench.py may load it.
//
// FP64 CPU restatement of the rigid-body algorithms that SAPIEN's PinocchioModel
// delegates to Pinocchio v2.6.18 (the third-party dependency is not vendored in
// the reference; its call sites are pinocchio/pinocchio_model.cpp:42-297).
// The formulation here is deliberately the *textbook* one (Featherstone, "Rigid
// Body Dynamics Algorithms", 2008): local joint frames, explicit 6x6 Plucker
// matrices, spatial inertia matrices, and Featherstone's tables for RNEA (5.1),
// CRBA (6.2) and ABA (7.1). The product kernels use a different, world-frame
// vector formulation, so a shared misconception has to survive two derivations.
//
// Spatial conventions follow Pinocchio (and therefore the reference outputs):
// motion = [linear; angular], force = [linear; angular], SE3 aMb maps frame-b
// coordinates to frame a: x_a = R x_b + p. Jacobian rows are [linear; angular];
// the linear rows of a WORLD Jacobian are the velocity of the point at the world
// origin (pinocchio_model.cpp:180 getJointJacobian(..., WORLD)).
//
// Everything is templated on the scalar so the same code runs in double (the
// oracle), float (CPU FP32 twin for error budgeting), a forward-mode dual number
// (exact derivatives: Christoffel Coriolis matrix, gradient references) and an
// op-counting scalar (exact algorithmic FLOPs).
#pragma once
#include <array>
#include <cmath>
#include <cstdint>
#include <stdexcept>
#include <vector>

namespace oracle {

// ---------------------------------------------------------------- scalar helpers
inline double sval(double x) { return x; }
inline float sval(float x) { return x; }
using std::atan2;
using std::cos;
using std::sin;
using std::sqrt;
using std::acos;
using std::fabs;

// Forward-mode dual number: value + one directional derivative.
struct Dual {
  double v = 0, d = 0;
  Dual() = default;
  Dual(double x) : v(x), d(0) {}
  Dual(double x, double dx) : v(x), d(dx) {}
};
inline Dual operator+(Dual a, Dual b) { return {a.v + b.v, a.d + b.d}; }
inline Dual operator-(Dual a, Dual b) { return {a.v - b.v, a.d - b.d}; }
inline Dual operator-(Dual a) { return {-a.v, -a.d}; }
inline Dual operator*(Dual a, Dual b) { return {a.v * b.v, a.d * b.v + a.v * b.d}; }
inline Dual operator/(Dual a, Dual b) { return {a.v / b.v, (a.d * b.v - a.v * b.d) / (b.v * b.v)}; }
inline Dual &operator+=(Dual &a, Dual b) { return a = a + b; }
inline Dual &operator-=(Dual &a, Dual b) { return a = a - b; }
inline Dual &operator*=(Dual &a, Dual b) { return a = a * b; }
inline Dual &operator/=(Dual &a, Dual b) { return a = a / b; }
inline bool operator<(Dual a, Dual b) { return a.v < b.v; }
inline bool operator>(Dual a, Dual b) { return a.v > b.v; }
inline bool operator<=(Dual a, Dual b) { return a.v <= b.v; }
inline bool operator>=(Dual a, Dual b) { return a.v >= b.v; }
inline Dual sin(Dual a) { return {std::sin(a.v), a.d * std::cos(a.v)}; }
inline Dual cos(Dual a) { return {std::cos(a.v), -a.d * std::sin(a.v)}; }
inline Dual sqrt(Dual a) {
  double s = std::sqrt(a.v);
  return {s, s > 0 ? a.d / (2 * s) : 0.0};
}
inline Dual atan2(Dual y, Dual x) {
  double r2 = x.v * x.v + y.v * y.v;
  return {std::atan2(y.v, x.v), (x.v * y.d - y.v * x.d) / r2};
}
inline Dual acos(Dual a) { return {std::acos(a.v), -a.d / std::sqrt(1 - a.v * a.v)}; }
inline Dual fabs(Dual a) { return a.v < 0 ? -a : a; }
inline double sval(Dual x) { return x.v; }

// Op-counting scalar: counts every arithmetic operation executed on it.
struct OpCount {
  double v = 0;
  static inline uint64_t adds = 0, muls = 0, divs = 0, sqrts = 0, trig = 0;
  OpCount() = default;
  OpCount(double x) : v(x) {}
  static void reset() { adds = muls = divs = sqrts = trig = 0; }
};
inline OpCount operator+(OpCount a, OpCount b) { ++OpCount::adds; return {a.v + b.v}; }
inline OpCount operator-(OpCount a, OpCount b) { ++OpCount::adds; return {a.v - b.v}; }
inline OpCount operator-(OpCount a) { return {-a.v}; }
inline OpCount operator*(OpCount a, OpCount b) { ++OpCount::muls; return {a.v * b.v}; }
inline OpCount operator/(OpCount a, OpCount b) { ++OpCount::divs; return {a.v / b.v}; }
inline OpCount &operator+=(OpCount &a, OpCount b) { return a = a + b; }
inline OpCount &operator-=(OpCount &a, OpCount b) { return a = a - b; }
inline OpCount &operator*=(OpCount &a, OpCount b) { return a = a * b; }
inline OpCount &operator/=(OpCount &a, OpCount b) { return a = a / b; }
inline bool operator<(OpCount a, OpCount b) { return a.v < b.v; }
inline bool operator>(OpCount a, OpCount b) { return a.v > b.v; }
inline bool operator<=(OpCount a, OpCount b) { return a.v <= b.v; }
inline bool operator>=(OpCount a, OpCount b) { return a.v >= b.v; }
inline OpCount sin(OpCount a) { ++OpCount::trig; return {std::sin(a.v)}; }
inline OpCount cos(OpCount a) { ++OpCount::trig; return {std::cos(a.v)}; }
inline OpCount sqrt(OpCount a) { ++OpCount::sqrts; return {std::sqrt(a.v)}; }
inline OpCount atan2(OpCount y, OpCount x) { ++OpCount::trig; return {std::atan2(y.v, x.v)}; }
inline OpCount acos(OpCount a) { ++OpCount::trig; return {std::acos(a.v)}; }
inline OpCount fabs(OpCount a) { return {std::fabs(a.v)}; }
inline double sval(OpCount x) { return x.v; }

// ---------------------------------------------------------------- small matrices
template <class S, int R, int C> struct Mat {
  S a[R][C];
  Mat() {
    for (int i = 0; i < R; ++i)
      for (int j = 0; j < C; ++j) a[i][j] = S(0.0);
  }
  S &operator()(int i, int j) { return a[i][j]; }
  const S &operator()(int i, int j) const { return a[i][j]; }
  static Mat identity() {
    Mat m;
    for (int i = 0; i < R && i < C; ++i) m.a[i][i] = S(1.0);
    return m;
  }
};
template <class S, int N> using Vec = Mat<S, N, 1>;
template <class S> using M3 = Mat<S, 3, 3>;
template <class S> using V3 = Vec<S, 3>;
template <class S> using M6 = Mat<S, 6, 6>;
template <class S> using V6 = Vec<S, 6>;

template <class S, int R, int K, int C>
Mat<S, R, C> operator*(const Mat<S, R, K> &A, const Mat<S, K, C> &B) {
  Mat<S, R, C> out;
  for (int i = 0; i < R; ++i)
    for (int j = 0; j < C; ++j) {
      S acc(0.0);
      for (int k = 0; k < K; ++k) acc += A.a[i][k] * B.a[k][j];
      out.a[i][j] = acc;
    }
  return out;
}
template <class S, int R, int C>
Mat<S, R, C> operator+(const Mat<S, R, C> &A, const Mat<S, R, C> &B) {
  Mat<S, R, C> o;
  for (int i = 0; i < R; ++i)
    for (int j = 0; j < C; ++j) o.a[i][j] = A.a[i][j] + B.a[i][j];
  return o;
}
template <class S, int R, int C>
Mat<S, R, C> operator-(const Mat<S, R, C> &A, const Mat<S, R, C> &B) {
  Mat<S, R, C> o;
  for (int i = 0; i < R; ++i)
    for (int j = 0; j < C; ++j) o.a[i][j] = A.a[i][j] - B.a[i][j];
  return o;
}
template <class S, int R, int C> Mat<S, R, C> scale(const Mat<S, R, C> &A, S s) {
  Mat<S, R, C> o;
  for (int i = 0; i < R; ++i)
    for (int j = 0; j < C; ++j) o.a[i][j] = A.a[i][j] * s;
  return o;
}
template <class S, int R, int C> Mat<S, C, R> tr(const Mat<S, R, C> &A) {
  Mat<S, C, R> o;
  for (int i = 0; i < R; ++i)
    for (int j = 0; j < C; ++j) o.a[j][i] = A.a[i][j];
  return o;
}
template <class S> M3<S> skew(const V3<S> &v) {
  M3<S> m;
  m(0, 1) = -v(2, 0); m(0, 2) = v(1, 0);
  m(1, 0) = v(2, 0);  m(1, 2) = -v(0, 0);
  m(2, 0) = -v(1, 0); m(2, 1) = v(0, 0);
  return m;
}
template <class S> V3<S> v3(S x, S y, S z) {
  V3<S> v;
  v(0, 0) = x; v(1, 0) = y; v(2, 0) = z;
  return v;
}
template <class S> V3<S> cross(const V3<S> &a, const V3<S> &b) { return skew(a) * b; }
template <class S> S dot3(const V3<S> &a, const V3<S> &b) {
  return a(0, 0) * b(0, 0) + a(1, 0) * b(1, 0) + a(2, 0) * b(2, 0);
}
template <class S, int N> S dotN(const Vec<S, N> &a, const Vec<S, N> &b) {
  S acc(0.0);
  for (int i = 0; i < N; ++i) acc += a(i, 0) * b(i, 0);
  return acc;
}

// 6x6 assembly from 3x3 blocks [[A, B], [C, D]].
template <class S> M6<S> blocks(const M3<S> &A, const M3<S> &B, const M3<S> &C, const M3<S> &D) {
  M6<S> m;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      m(i, j) = A(i, j); m(i, j + 3) = B(i, j);
      m(i + 3, j) = C(i, j); m(i + 3, j + 3) = D(i, j);
    }
  return m;
}
template <class S> V6<S> v6(const V3<S> &lin, const V3<S> &ang) {
  V6<S> o;
  for (int i = 0; i < 3; ++i) { o(i, 0) = lin(i, 0); o(i + 3, 0) = ang(i, 0); }
  return o;
}

// ---------------------------------------------------------------- rigid transforms
template <class S> struct SE3 {
  M3<S> R = M3<S>::identity();
  V3<S> p;
};
template <class S> SE3<S> compose(const SE3<S> &A, const SE3<S> &B) {
  SE3<S> o;
  o.R = A.R * B.R;
  o.p = A.R * B.p + A.p;
  return o;
}
template <class S> SE3<S> inverse(const SE3<S> &A) {
  SE3<S> o;
  o.R = tr(A.R);
  o.p = scale(o.R * A.p, S(-1.0));
  return o;
}
// Plucker motion transform of aMb (b -> a): [[R, [p]x R], [0, R]].
template <class S> M6<S> motionX(const SE3<S> &M) {
  M3<S> Z;
  return blocks(M.R, skew(M.p) * M.R, Z, M.R);
}
// Inverse motion transform (a -> b): [[R^T, -R^T [p]x], [0, R^T]].
template <class S> M6<S> motionXinv(const SE3<S> &M) {
  M3<S> Rt = tr(M.R), Z;
  return blocks(Rt, scale(Rt * skew(M.p), S(-1.0)), Z, Rt);
}
// Motion cross-product matrix crm(m): m x n, for m = [v; w].
template <class S> M6<S> crm(const V6<S> &m) {
  V3<S> v = v3<S>(m(0, 0), m(1, 0), m(2, 0)), w = v3<S>(m(3, 0), m(4, 0), m(5, 0));
  M3<S> Z;
  return blocks(skew(w), skew(v), Z, skew(w));
}
// Force cross-product matrix crf(m) = -crm(m)^T.
template <class S> M6<S> crf(const V6<S> &m) { return scale(tr(crm(m)), S(-1.0)); }

// Rodrigues rotation about a unit axis given (cos, sin) of the angle.
template <class S> M3<S> axis_rotation(const V3<S> &a, S c, S s) {
  M3<S> K = skew(a);
  return M3<S>::identity() + scale(K, s) + scale(K * K, S(1.0) - c);
}

// ---------------------------------------------------------------- model
enum JointType : int32_t { REVOLUTE = 0, PRISMATIC = 1, CONTINUOUS = 2 };

// Model tables in Pinocchio's internal (topological, depth-first) joint order.
// Built by oracle/urdf_model.py from URDF text with Pinocchio v2.6.18's URDF
// semantics (fixed joints merged into the parent body, inertial origin rotates
// the inertia tensor, link frames attached to their parent joint).
struct Model {
  int nj = 0;                         // movable joints, universe excluded
  std::vector<int> parent;            // parent joint, -1 = universe
  std::vector<int> jtype;             // JointType
  std::vector<std::array<double, 3>> axis;      // unit axis in the joint frame
  std::vector<std::array<double, 12>> jplace;   // parent-joint frame -> joint frame, R (row-major) | p
  std::vector<std::array<double, 10>> inertia;  // m, com[3], I_com[xx xy xz yy yz zz], joint frame
  std::vector<int> vuser;             // user (SAPIEN-order) dof index of the joint
  int nlinks = 0;
  std::vector<int> link_joint;        // joint the link's frame is attached to, -1 = universe
  std::vector<std::array<double, 12>> link_place;  // joint frame -> link frame
  double gravity[3] = {0, 0, -9.81};
};

template <class S> SE3<S> se3_from(const std::array<double, 12> &a) {
  SE3<S> M;
  for (int i = 0; i < 3; ++i) {
    for (int j = 0; j < 3; ++j) M.R(i, j) = S(a[3 * i + j]);
    M.p(i, 0) = S(a[9 + i]);
  }
  return M;
}
template <class S> V3<S> axis_of(const Model &m, int j) {
  return v3<S>(S(m.axis[j][0]), S(m.axis[j][1]), S(m.axis[j][2]));
}
// Joint motion subspace S_j in its own frame: revolute [0; a], prismatic [a; 0].
template <class S> V6<S> motion_subspace(const Model &m, int j) {
  V3<S> a = axis_of<S>(m, j), z;
  return m.jtype[j] == PRISMATIC ? v6(a, z) : v6(z, a);
}
// Spatial inertia matrix of the joint's (merged) body in the joint frame.
template <class S> M6<S> spatial_inertia(const Model &m, int j) {
  const auto &I = m.inertia[j];
  S mass(I[0]);
  V3<S> c = v3<S>(S(I[1]), S(I[2]), S(I[3]));
  M3<S> Ic;
  Ic(0, 0) = S(I[4]); Ic(0, 1) = S(I[5]); Ic(0, 2) = S(I[6]);
  Ic(1, 0) = S(I[5]); Ic(1, 1) = S(I[7]); Ic(1, 2) = S(I[8]);
  Ic(2, 0) = S(I[6]); Ic(2, 1) = S(I[8]); Ic(2, 2) = S(I[9]);
  M3<S> C = skew(c);
  M3<S> mI = scale(M3<S>::identity(), mass);
  return blocks(mI, scale(C, -mass), scale(C, mass), Ic - scale(C * C, mass));
}
// Joint transform X_J(q) for the joint's own dof value (user q).
template <class S> SE3<S> joint_transform(const Model &m, int j, S q) {
  SE3<S> M;
  V3<S> a = axis_of<S>(m, j);
  if (m.jtype[j] == PRISMATIC) {
    M.p = scale(a, q);
  } else {
    M.R = axis_rotation(a, cos(q), sin(q));
  }
  return M;
}

// ---------------------------------------------------------------- kinematics
// forwardKinematics (pinocchio_model.cpp:146-148): oMi for every joint.
template <class S> void forward_kinematics(const Model &m, const S *q, std::vector<SE3<S>> &oMi) {
  oMi.resize(m.nj);
  for (int j = 0; j < m.nj; ++j) {
    SE3<S> liMi = compose(se3_from<S>(m.jplace[j]), joint_transform(m, j, q[m.vuser[j]]));
    oMi[j] = m.parent[j] < 0 ? liMi : compose(oMi[m.parent[j]], liMi);
  }
}
// getLinkPose (pinocchio_model.cpp:150-161): joint2world * link2joint.
template <class S> SE3<S> link_pose(const Model &m, const std::vector<SE3<S>> &oMi, int l) {
  SE3<S> l2j = se3_from<S>(m.link_place[l]);
  int j = m.link_joint[l];
  return j < 0 ? l2j : compose(oMi[j], l2j);
}
// Quaternion (w, x, y, z) of a rotation matrix by Shoemake's method (pick the
// largest of w, x, y, z to divide by). Same method as Eigen::Quaterniond(R)
// used at pinocchio_model.cpp:159.
template <class S> void quat_from_matrix(const M3<S> &R, S out[4]) {
  S t = R(0, 0) + R(1, 1) + R(2, 2);
  if (t > S(0.0)) {
    S r = sqrt(t + S(1.0));
    S h = S(0.5) / r;
    out[0] = S(0.5) * r;
    out[1] = (R(2, 1) - R(1, 2)) * h;
    out[2] = (R(0, 2) - R(2, 0)) * h;
    out[3] = (R(1, 0) - R(0, 1)) * h;
  } else {
    int i = 0;
    if (R(1, 1) > R(0, 0)) i = 1;
    if (R(2, 2) > R(i, i)) i = 2;
    int j = (i + 1) % 3, k = (j + 1) % 3;
    S r = sqrt(R(i, i) - R(j, j) - R(k, k) + S(1.0));
    S h = S(0.5) / r;
    S xyz[3];
    xyz[i] = S(0.5) * r;
    xyz[j] = (R(j, i) + R(i, j)) * h;
    xyz[k] = (R(k, i) + R(i, k)) * h;
    out[0] = (R(k, j) - R(j, k)) * h;
    out[1] = xyz[0]; out[2] = xyz[1]; out[3] = xyz[2];
  }
}

// is joint a an ancestor-or-self of joint b?
inline bool supports(const Model &m, int a, int b) {
  for (int j = b; j >= 0; j = m.parent[j])
    if (j == a) return true;
  return false;
}

// computeJointJacobians (pinocchio_model.cpp:163-165): column of joint k is
// oMi[k].act(S_k) in the world frame; stored in user column order.
template <class S>
void world_jacobian(const Model &m, const std::vector<SE3<S>> &oMi, std::vector<V6<S>> &cols) {
  cols.resize(m.nj);
  for (int j = 0; j < m.nj; ++j) cols[j] = motionX(oMi[j]) * motion_subspace<S>(m, j);
}

// getLinkJacobian (pinocchio_model.cpp:167-186): support columns of the link's
// parent joint, zero elsewhere; local=True maps by Ad(link2world)^-1 (body frame).
// out is 6 x nv row-major in user column order.
template <class S>
void link_jacobian(const Model &m, const std::vector<SE3<S>> &oMi, const std::vector<V6<S>> &cols,
                   int l, bool local, S *out) {
  int nv = m.nj;
  for (int i = 0; i < 6 * nv; ++i) out[i] = S(0.0);
  int pj = m.link_joint[l];
  SE3<S> l2w = link_pose(m, oMi, l);
  M6<S> Ainv = motionXinv(l2w);
  for (int k = 0; k < m.nj; ++k) {
    if (pj < 0 || !supports(m, k, pj)) continue;
    V6<S> c = local ? Ainv * cols[k] : cols[k];
    for (int r = 0; r < 6; ++r) out[r * nv + m.vuser[k]] = c(r, 0);
  }
}

// ---------------------------------------------------------------- dynamics
// rnea (pinocchio_model.cpp:214-219), Featherstone RBDA Table 5.1, local frames.
template <class S>
void rnea(const Model &m, const S *q, const S *qd, const S *qdd, S *tau) {
  int n = m.nj;
  std::vector<M6<S>> Xup(n);
  std::vector<V6<S>> v(n), a(n), f(n), Sj(n);
  V6<S> a0;
  for (int i = 0; i < 3; ++i) a0(i, 0) = S(-m.gravity[i]);
  for (int j = 0; j < n; ++j) {
    int u = m.vuser[j];
    SE3<S> liMi = compose(se3_from<S>(m.jplace[j]), joint_transform(m, j, q[u]));
    Xup[j] = motionXinv(liMi);
    Sj[j] = motion_subspace<S>(m, j);
    V6<S> vJ = scale(Sj[j], qd[u]);
    int p = m.parent[j];
    V6<S> vp, ap = a0;
    if (p >= 0) { vp = v[p]; ap = a[p]; }
    v[j] = Xup[j] * vp + vJ;
    a[j] = Xup[j] * ap + scale(Sj[j], qdd[u]) + crm(v[j]) * vJ;
    M6<S> I = spatial_inertia<S>(m, j);
    f[j] = I * a[j] + crf(v[j]) * (I * v[j]);
  }
  for (int j = n - 1; j >= 0; --j) {
    tau[m.vuser[j]] = dotN(Sj[j], f[j]);
    int p = m.parent[j];
    if (p >= 0) f[p] = f[p] + tr(Xup[j]) * f[j];
  }
}

// crba (pinocchio_model.cpp:201-206), Featherstone RBDA Table 6.2; the full
// symmetric matrix in user order (the reference symmetrizes at :203-204).
template <class S> void crba(const Model &m, const S *q, S *M) {
  int n = m.nj;
  std::vector<M6<S>> Xup(n), Ic(n);
  std::vector<V6<S>> Sj(n);
  for (int j = 0; j < n; ++j) {
    SE3<S> liMi = compose(se3_from<S>(m.jplace[j]), joint_transform(m, j, q[m.vuser[j]]));
    Xup[j] = motionXinv(liMi);
    Sj[j] = motion_subspace<S>(m, j);
    Ic[j] = spatial_inertia<S>(m, j);
  }
  for (int j = n - 1; j >= 0; --j) {
    int p = m.parent[j];
    if (p >= 0) Ic[p] = Ic[p] + tr(Xup[j]) * Ic[j] * Xup[j];
  }
  for (int i = 0; i < n * n; ++i) M[i] = S(0.0);
  for (int i = 0; i < n; ++i) {
    V6<S> F = Ic[i] * Sj[i];
    int ui = m.vuser[i];
    M[ui * n + ui] = dotN(Sj[i], F);
    int j = i;
    while (m.parent[j] >= 0) {
      F = tr(Xup[j]) * F;
      j = m.parent[j];
      int uj = m.vuser[j];
      M[ui * n + uj] = M[uj * n + ui] = dotN(Sj[j], F);
    }
  }
}

// aba (pinocchio_model.cpp:221-226), Featherstone RBDA Table 7.1.
template <class S>
void aba(const Model &m, const S *q, const S *qd, const S *tau, S *qdd) {
  int n = m.nj;
  std::vector<M6<S>> Xup(n), IA(n);
  std::vector<V6<S>> Sj(n), v(n), c(n), pA(n), U(n), a(n);
  std::vector<S> D(n), u(n);
  for (int j = 0; j < n; ++j) {
    int uj = m.vuser[j];
    SE3<S> liMi = compose(se3_from<S>(m.jplace[j]), joint_transform(m, j, q[uj]));
    Xup[j] = motionXinv(liMi);
    Sj[j] = motion_subspace<S>(m, j);
    V6<S> vJ = scale(Sj[j], qd[uj]);
    int p = m.parent[j];
    V6<S> vp;
    if (p >= 0) vp = v[p];
    v[j] = Xup[j] * vp + vJ;
    c[j] = crm(v[j]) * vJ;
    IA[j] = spatial_inertia<S>(m, j);
    pA[j] = crf(v[j]) * (IA[j] * v[j]);
  }
  for (int j = n - 1; j >= 0; --j) {
    U[j] = IA[j] * Sj[j];
    D[j] = dotN(Sj[j], U[j]);
    u[j] = tau[m.vuser[j]] - dotN(Sj[j], pA[j]);
    int p = m.parent[j];
    if (p >= 0) {
      M6<S> Ia = IA[j] - scale(U[j] * tr(U[j]), S(1.0) / D[j]);
      V6<S> pa = pA[j] + Ia * c[j] + scale(U[j], u[j] / D[j]);
      IA[p] = IA[p] + tr(Xup[j]) * Ia * Xup[j];
      pA[p] = pA[p] + tr(Xup[j]) * pa;
    }
  }
  V6<S> a0;
  for (int i = 0; i < 3; ++i) a0(i, 0) = S(-m.gravity[i]);
  for (int j = 0; j < n; ++j) {
    int p = m.parent[j];
    V6<S> ap = p >= 0 ? a[p] : a0;
    a[j] = Xup[j] * ap + c[j];
    S qddj = (u[j] - dotN(U[j], a[j])) / D[j];
    qdd[m.vuser[j]] = qddj;
    a[j] = a[j] + scale(Sj[j], qddj);
  }
}

// ---------------------------------------------------------------- SE(3) log and its Jacobian
// log3 of a rotation (Pinocchio's explog restated from the mathematics:
// theta = acos((tr R - 1) / 2), axis from the skew part, and near theta = pi the
// axis magnitudes from the diagonal R_ii = cos + (1 - cos) k_i^2).
template <class S> V3<S> log3(const M3<S> &R, S &theta) {
  S t = R(0, 0) + R(1, 1) + R(2, 2);
  const double PI = 3.14159265358979323846;
  if (t >= S(3.0)) theta = S(0.0);
  else if (t <= S(-1.0)) theta = S(PI);
  else theta = acos((t - S(1.0)) / S(2.0));
  V3<S> w;
  if (sval(theta) > PI - 1e-2) {
    S c = cos(theta);
    S k = theta * theta / (S(1.0) - c);
    S sgn[3] = {R(2, 1) > R(1, 2) ? S(1.0) : S(-1.0), R(0, 2) > R(2, 0) ? S(1.0) : S(-1.0),
                R(1, 0) > R(0, 1) ? S(1.0) : S(-1.0)};
    for (int i = 0; i < 3; ++i) {
      S x = (R(i, i) - c) * k;
      w(i, 0) = sval(x) > 0 ? sgn[i] * sqrt(x) : S(0.0);
    }
  } else {
    // factor 1/2 below theta = eps^(1/4) = 2^-13, as Pinocchio 2.6's log3
    S h = sval(theta) > 1.220703125e-4 ? theta / sin(theta) / S(2.0) : S(0.5);
    w = v3<S>(h * (R(2, 1) - R(1, 2)), h * (R(0, 2) - R(2, 0)), h * (R(1, 0) - R(0, 1)));
  }
  return w;
}
// Coefficients of the SE(3) log / its inverse right Jacobian as functions of the
// rotation angle t: alpha = (t/2)cot(t/2), beta = (1 - alpha)/t^2 and
// dbeta = (d beta / dt) / t, with their Taylor expansions near zero.
template <class S> void log_coeffs(S t, S &alpha, S &beta, S &dbeta) {
  S t2 = t * t;
  if (sval(t) < 1.220703125e-4) {  // series below eps^(1/4) = 2^-13, as Pinocchio 2.6's log6
    alpha = S(1.0) - t2 / S(12.0) - t2 * t2 / S(720.0);
    beta = S(1.0) / S(12.0) + t2 / S(720.0);
    dbeta = S(1.0) / S(360.0);
  } else {
    S st = sin(t), ct = cos(t);
    alpha = t * st / (S(2.0) * (S(1.0) - ct));
    beta = S(1.0) / t2 - st / (S(2.0) * t * (S(1.0) - ct));
    dbeta = S(-2.0) / (t2 * t2) + (S(1.0) + st / t) / (S(2.0) * t2 * (S(1.0) - ct));
  }
}
// log6(M) = [V(w)^-1 p; w] with V^-1 = alpha I - [w]x/2 + beta w w^T.
template <class S> V6<S> log6(const SE3<S> &M) {
  S t;
  V3<S> w = log3(M.R, t);
  S alpha, beta, dbeta;
  log_coeffs(t, alpha, beta, dbeta);
  V3<S> v = scale(M.p, alpha) - scale(cross(w, M.p), S(0.5)) + scale(w, beta * dot3(w, M.p));
  return v6(v, w);
}
// Jlog6(M) = d log6(M exp(d)) / dd at d = 0: [[A, C A], [0, A]] with
// A = Jr^-1(w) = alpha I + beta w w^T + [w]x/2 and C = d(V^-1(w) p)/dw.
template <class S> M6<S> Jlog6(const SE3<S> &M) {
  S t;
  V3<S> w = log3(M.R, t);
  S alpha, beta, dbeta;
  log_coeffs(t, alpha, beta, dbeta);
  const V3<S> &p = M.p;
  M3<S> wwT = w * tr(w);
  M3<S> A = scale(M3<S>::identity(), alpha) + scale(wwT, beta) + scale(skew(w), S(0.5));
  S wp = dot3(w, p);
  M3<S> C = scale(wwT, dbeta * wp) - scale(p * tr(w), t * t * dbeta + S(2.0) * beta) +
            scale(w * tr(p), beta) + scale(M3<S>::identity(), wp * beta) + scale(skew(p), S(0.5));
  M3<S> Z;
  return blocks(A, C * A, Z, A);
}

}  // namespace oracle
