// World-frame spatial algebra for the kernels (device, templated on float/double).
//
// Motion vectors are [linear; angular] with the linear part taken at the world
// origin, forces [linear; angular] likewise (the reference's convention,
// pinocchio_model.cpp:180). In this frame a joint's Jacobian column is fixed in
// its body, so velocities and accelerations are prefix sums over ancestors and
// forces are subtree sums: no per-joint change of frame is needed after FK.
#pragma once
#include <hip/hip_runtime.h>

#include "tables.h"

namespace sapamd {

template <class T> struct V3 {
  T x, y, z;
};
template <class T> __device__ __forceinline__ V3<T> mk(T x, T y, T z) { return {x, y, z}; }
template <class T> __device__ __forceinline__ V3<T> operator+(V3<T> a, V3<T> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <class T> __device__ __forceinline__ V3<T> operator-(V3<T> a, V3<T> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <class T> __device__ __forceinline__ V3<T> operator*(T s, V3<T> a) { return {s * a.x, s * a.y, s * a.z}; }
template <class T> __device__ __forceinline__ T dot(V3<T> a, V3<T> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <class T> __device__ __forceinline__ V3<T> cross(V3<T> a, V3<T> b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}

// Row-major 3x3.
template <class T> struct M3 {
  T a[9];
};
template <class T> __device__ __forceinline__ V3<T> mv(const T *R, V3<T> v) {
  return {R[0] * v.x + R[1] * v.y + R[2] * v.z, R[3] * v.x + R[4] * v.y + R[5] * v.z,
          R[6] * v.x + R[7] * v.y + R[8] * v.z};
}
template <class T> __device__ __forceinline__ V3<T> mtv(const T *R, V3<T> v) {
  return {R[0] * v.x + R[3] * v.y + R[6] * v.z, R[1] * v.x + R[4] * v.y + R[7] * v.z,
          R[2] * v.x + R[5] * v.y + R[8] * v.z};
}
template <class T> __device__ __forceinline__ void mm(const T *A, const T *B, T *C) {
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) C[3 * i + j] = A[3 * i] * B[j] + A[3 * i + 1] * B[3 + j] + A[3 * i + 2] * B[6 + j];
}
template <class T> __device__ __forceinline__ V3<T> load3(const T *p) { return {p[0], p[1], p[2]}; }

// Motion / force 6-vectors.
template <class T> struct S6 {
  V3<T> l, a;
};
template <class T> __device__ __forceinline__ S6<T> operator+(S6<T> x, S6<T> y) { return {x.l + y.l, x.a + y.a}; }
template <class T> __device__ __forceinline__ S6<T> operator-(S6<T> x, S6<T> y) { return {x.l - y.l, x.a - y.a}; }
template <class T> __device__ __forceinline__ S6<T> operator*(T s, S6<T> x) { return {s * x.l, s * x.a}; }
template <class T> __device__ __forceinline__ T dot6(S6<T> x, S6<T> y) { return dot(x.l, y.l) + dot(x.a, y.a); }
template <class T> __device__ __forceinline__ S6<T> zero6() { return {{T(0), T(0), T(0)}, {T(0), T(0), T(0)}}; }
// motion x motion: [w1 x v2 + v1 x w2; w1 x w2]
template <class T> __device__ __forceinline__ S6<T> crossm(S6<T> m1, S6<T> m2) {
  return {cross(m1.a, m2.l) + cross(m1.l, m2.a), cross(m1.a, m2.a)};
}
// motion x* force: [w x f; w x n + v x f]
template <class T> __device__ __forceinline__ S6<T> crossf(S6<T> m, S6<T> f) {
  return {cross(m.a, f.l), cross(m.a, f.a) + cross(m.l, f.l)};
}

// Rigid-body inertia in the world frame: mass, COM, rotational inertia about the
// COM (symmetric: xx xy xz yy yz zz).
template <class T> struct WInertia {
  T m;
  V3<T> c;
  T I[6];
};
template <class T> __device__ __forceinline__ V3<T> sym_mv(const T *I, V3<T> w) {
  return {I[0] * w.x + I[1] * w.y + I[2] * w.z, I[1] * w.x + I[3] * w.y + I[4] * w.z,
          I[2] * w.x + I[4] * w.y + I[5] * w.z};
}
// Momentum h = I V: l = m (v + w x c), k = I_c w + c x l.
template <class T> __device__ __forceinline__ S6<T> inertia_apply(const WInertia<T> &Y, S6<T> V) {
  V3<T> l = Y.m * (V.l + cross(V.a, Y.c));
  return {l, sym_mv(Y.I, V.a) + cross(Y.c, l)};
}
// World inertia of a joint's body: c_w = R c + p, I_w = R I_c R^T.
template <class T>
__device__ __forceinline__ WInertia<T> world_inertia(const JointRec<T> &J, const T *R, V3<T> p) {
  WInertia<T> Y;
  Y.m = J.m;
  Y.c = mv(R, load3(J.c)) + p;
  const T *I = J.Ic;
  // RI = R * Ic (Ic symmetric)
  T RI[9];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    RI[3 * i + 0] = R[3 * i] * I[0] + R[3 * i + 1] * I[1] + R[3 * i + 2] * I[2];
    RI[3 * i + 1] = R[3 * i] * I[1] + R[3 * i + 1] * I[3] + R[3 * i + 2] * I[4];
    RI[3 * i + 2] = R[3 * i] * I[2] + R[3 * i + 1] * I[4] + R[3 * i + 2] * I[5];
  }
  Y.I[0] = RI[0] * R[0] + RI[1] * R[1] + RI[2] * R[2];
  Y.I[1] = RI[0] * R[3] + RI[1] * R[4] + RI[2] * R[5];
  Y.I[2] = RI[0] * R[6] + RI[1] * R[7] + RI[2] * R[8];
  Y.I[3] = RI[3] * R[3] + RI[4] * R[4] + RI[5] * R[5];
  Y.I[4] = RI[3] * R[6] + RI[4] * R[7] + RI[5] * R[8];
  Y.I[5] = RI[6] * R[6] + RI[7] * R[7] + RI[8] * R[8];
  return Y;
}

// Additive spatial inertia about the world origin (sums over subtrees):
// mass, first moment h = m c, I_O = I_c + m (|c|^2 I - c c^T).
template <class T> struct AInertia {
  T m;
  V3<T> h;
  T I[6];
};
template <class T> __device__ __forceinline__ AInertia<T> additive(const WInertia<T> &Y) {
  AInertia<T> A;
  A.m = Y.m;
  A.h = Y.m * Y.c;
  T cc = dot(Y.c, Y.c);
  A.I[0] = Y.I[0] + Y.m * (cc - Y.c.x * Y.c.x);
  A.I[1] = Y.I[1] - Y.m * Y.c.x * Y.c.y;
  A.I[2] = Y.I[2] - Y.m * Y.c.x * Y.c.z;
  A.I[3] = Y.I[3] + Y.m * (cc - Y.c.y * Y.c.y);
  A.I[4] = Y.I[4] - Y.m * Y.c.y * Y.c.z;
  A.I[5] = Y.I[5] + Y.m * (cc - Y.c.z * Y.c.z);
  return A;
}
template <class T> __device__ __forceinline__ void add_to(AInertia<T> &a, const AInertia<T> &b) {
  a.m += b.m;
  a.h = a.h + b.h;
#pragma unroll
  for (int i = 0; i < 6; ++i) a.I[i] += b.I[i];
}
// Y V = [m v + w x h; I_O w + h x v]
template <class T> __device__ __forceinline__ S6<T> apply(const AInertia<T> &Y, S6<T> V) {
  return {Y.m * V.l + cross(V.a, Y.h), sym_mv(Y.I, V.a) + cross(Y.h, V.l)};
}

// Local joint transform and world pose of joint j given its parent's world pose.
template <class T>
__device__ __forceinline__ void joint_world(const JointRec<T> &J, T q, T s, T c, const T *Rp, V3<T> pp, bool root,
                                            T *R, V3<T> &p) {
  T Rl[9];
  V3<T> pl = load3(J.p);
  if (J.kind == KIND_PRISMATIC) {
#pragma unroll
    for (int k = 0; k < 9; ++k) Rl[k] = J.A[k];
    pl = pl + q * load3(J.ap);
  } else {
    T omc = T(1) - c;
#pragma unroll
    for (int k = 0; k < 9; ++k) Rl[k] = J.A[k] + s * J.Bq[k] + omc * J.Cq[k];
  }
  if (root) {
#pragma unroll
    for (int k = 0; k < 9; ++k) R[k] = Rl[k];
    p = pl;
  } else {
    mm(Rp, Rl, R);
    p = mv(Rp, pl) + pp;
  }
}

// World Jacobian column of a joint with world rotation R and origin p.
template <class T> __device__ __forceinline__ S6<T> jcol(const JointRec<T> &J, const T *R, V3<T> p) {
  V3<T> aw = mv(R, load3(J.axis));
  if (J.kind == KIND_PRISMATIC) return {aw, {T(0), T(0), T(0)}};
  return {cross(p, aw), aw};
}

template <class T> __device__ __forceinline__ void sincos_t(T x, T *s, T *c);
template <> __device__ __forceinline__ void sincos_t<float>(float x, float *s, float *c) { sincosf(x, s, c); }
template <> __device__ __forceinline__ void sincos_t<double>(double x, double *s, double *c) { sincos(x, s, c); }

// Quaternion (w, x, y, z) of a rotation matrix by Shoemake's method: divide by
// the largest of 4w^2, 4x^2, 4y^2, 4z^2 (w > 0 when the trace is positive).
template <class T> __device__ __forceinline__ void quat_of(const T *R, T *q) {
  T t = R[0] + R[4] + R[8];
  if (t > T(0)) {
    T r = sqrt(t + T(1));
    T h = T(0.5) / r;
    q[0] = T(0.5) * r;
    q[1] = (R[7] - R[5]) * h;
    q[2] = (R[2] - R[6]) * h;
    q[3] = (R[3] - R[1]) * h;
  } else if (R[0] >= R[4] && R[0] >= R[8]) {
    T r = sqrt(R[0] - R[4] - R[8] + T(1));
    T h = T(0.5) / r;
    q[1] = T(0.5) * r;
    q[2] = (R[3] + R[1]) * h;
    q[3] = (R[6] + R[2]) * h;
    q[0] = (R[7] - R[5]) * h;
  } else if (R[4] >= R[8]) {
    T r = sqrt(R[4] - R[8] - R[0] + T(1));
    T h = T(0.5) / r;
    q[2] = T(0.5) * r;
    q[3] = (R[7] + R[5]) * h;
    q[1] = (R[1] + R[3]) * h;
    q[0] = (R[2] - R[6]) * h;
  } else {
    T r = sqrt(R[8] - R[0] - R[4] + T(1));
    T h = T(0.5) / r;
    q[3] = T(0.5) * r;
    q[1] = (R[2] + R[6]) * h;
    q[2] = (R[5] + R[7]) * h;
    q[0] = (R[3] - R[1]) * h;
  }
}

}  // namespace sapamd
