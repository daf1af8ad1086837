// SE(3) helpers of the CLIK kernels (device, templated on float/double):
// body-frame images of motion vectors, the SE(3) log and its Jacobian, and the
// damped 6x6 solve, as used by compute_inverse_kinematics
// (pinocchio_model.cpp:228-300).
#pragma once
#include <hip/hip_runtime.h>

#include "spatial.h"

namespace sapamd {

// Body (local) frame image of a world motion vector under link pose (R, p):
// Ad(M)^-1 m = [R^T (l - p x a); R^T a].
template <class T> __device__ __forceinline__ S6<T> to_local(const T *R, V3<T> p, S6<T> m) {
  return {mtv(R, m.l - cross(p, m.a)), mtv(R, m.a)};
}

template <class T> __device__ __forceinline__ void vec_of(S6<T> x, T *v) {
  v[0] = x.l.x; v[1] = x.l.y; v[2] = x.l.z; v[3] = x.a.x; v[4] = x.a.y; v[5] = x.a.z;
}

// ---------------------------------------------------------------- SE(3) log for CLIK
// Rotation log in the form Pinocchio 2.6 evaluates it (the reference's
// log6 / Jlog6 inside compute_inverse_kinematics): theta = acos((tr R - 1) / 2)
// clipped to [0, pi]; below pi - 1e-2 the axial vector
// theta / (2 sin theta) (R - R^T)^v, with the factor 1 below theta = 2^-13
// (eps^(1/4)) in FP64 and 1 + theta^2 / 6 below 1e-3 in FP32; above, the axis
// magnitudes from the diagonal, R_ii = cos theta + (1 - cos theta) k_i^2, with
// the signs of the skew part. For a rotation any correct log agrees; this
// form also gives the reference's error vector when R is not orthonormal (a
// target built from a quaternion that is not unit length).
template <class T> __device__ __forceinline__ V3<T> log3_dev(const T *R, T &theta) {
  const T PI = T(3.14159265358979323846);
  const T tr = R[0] + R[4] + R[8];
  theta = tr >= T(3) ? T(0) : tr <= T(-1) ? PI : acos((tr - T(1)) * T(0.5));
  const V3<T> ax = {T(0.5) * (R[7] - R[5]), T(0.5) * (R[2] - R[6]), T(0.5) * (R[3] - R[1])};
  if (theta < PI - T(1e-2)) {
    const bool fp64 = sizeof(T) == 8;
    const T k = theta > (fp64 ? T(1.220703125e-4) : T(1e-3)) ? theta / sin(theta)
                : fp64 ? T(1) : T(1) + theta * theta / T(6);
    return k * ax;
  }
  const T c = cos(theta);
  const T b = theta * theta / (T(1) - c);
  const T x0 = (R[0] - c) * b, x1 = (R[4] - c) * b, x2 = (R[8] - c) * b;
  return {(R[7] > R[5] ? T(1) : T(-1)) * (x0 > T(0) ? sqrt(x0) : T(0)),
          (R[2] > R[6] ? T(1) : T(-1)) * (x1 > T(0) ? sqrt(x1) : T(0)),
          (R[3] > R[1] ? T(1) : T(-1)) * (x2 > T(0) ? sqrt(x2) : T(0))};
}

// alpha = (t/2) cot(t/2), beta = (1 - alpha) / t^2, dbeta = beta'(t) / t; series below 0.5.
template <class T> __device__ __forceinline__ void log_coeffs_dev(T t, T &alpha, T &beta, T &dbeta) {
  T t2 = t * t;
  if (t < T(0.5)) {
    alpha = T(1) - t2 * (T(1) / T(12) + t2 * (T(1) / T(720) + t2 * (T(1) / T(30240) + t2 / T(1209600))));
    beta = T(1) / T(12) + t2 * (T(1) / T(720) + t2 * (T(1) / T(30240) + t2 * (T(1) / T(1209600) + t2 / T(47900160))));
    dbeta = T(1) / T(360) + t2 * (T(1) / T(7560) + t2 * (T(1) / T(201600) + t2 / T(5987520)));
  } else {
    T st, ct;
    sincos_t(t, &st, &ct);
    // two reciprocals (1/t, 1/(1 - cos t)) instead of one division per term
    const T r = T(1) / t, u = T(1) / (T(1) - ct), r2 = r * r, hsu = T(0.5) * st * u;
    alpha = t * hsu;
    beta = r2 - r * hsu;
    dbeta = T(-2) * r2 * r2 + (T(1) + st * r) * (T(0.5) * u) * r2;
  }
}

template <class T>
__device__ __forceinline__ S6<T> log6_dev(const T *R, V3<T> p) {
  T t;
  V3<T> w = log3_dev(R, t);
  T al, be, db;
  log_coeffs_dev(t, al, be, db);
  return {al * p - T(0.5) * cross(w, p) + (be * dot(w, p)) * w, w};
}

// Jlog6 (6x6 row-major): [[A, C A], [0, A]], A = alpha I + beta w w^T + [w]x / 2,
// C = dbeta (w.p) w w^T - (t^2 dbeta + 2 beta) p w^T + beta w p^T + beta (w.p) I + [p]x / 2.
// jlog6_blocks gives the 3x3 blocks A and C; jlog6_core the assembled matrix.
template <class T>
__device__ __forceinline__ void jlog6_blocks(V3<T> w, T t, T al, T be, T db, V3<T> p, T *A, T *C) {
  T wv[3] = {w.x, w.y, w.z}, pv[3] = {p.x, p.y, p.z};
  T wp = dot(w, p);
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      A[3 * i + j] = be * wv[i] * wv[j] + (i == j ? al : T(0));
      C[3 * i + j] = db * wp * wv[i] * wv[j] - (t * t * db + T(2) * be) * pv[i] * wv[j] + be * wv[i] * pv[j] +
                     (i == j ? be * wp : T(0));
    }
  A[1] -= T(0.5) * w.z; A[2] += T(0.5) * w.y; A[5] -= T(0.5) * w.x;
  A[3] += T(0.5) * w.z; A[6] -= T(0.5) * w.y; A[7] += T(0.5) * w.x;
  C[1] -= T(0.5) * p.z; C[2] += T(0.5) * p.y; C[5] -= T(0.5) * p.x;
  C[3] += T(0.5) * p.z; C[6] -= T(0.5) * p.y; C[7] += T(0.5) * p.x;
}
template <class T>
__device__ __forceinline__ void jlog6_core(V3<T> w, T t, T al, T be, T db, V3<T> p, T *Jl) {
  T A[9], C[9];
  jlog6_blocks(w, t, al, be, db, p, A, C);
  T CA[9];
  mm(C, A, CA);
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      Jl[i * 6 + j] = A[3 * i + j];
      Jl[i * 6 + 3 + j] = CA[3 * i + j];
      Jl[(3 + i) * 6 + j] = T(0);
      Jl[(3 + i) * 6 + 3 + j] = A[3 * i + j];
    }
}

template <class T> __device__ __forceinline__ void jlog6_dev(const T *R, V3<T> p, T *Jl) {
  T t;
  V3<T> w = log3_dev(R, t);
  T al, be, db;
  log_coeffs_dev(t, al, be, db);
  jlog6_core(w, t, al, be, db, p, Jl);
}

// log6 of (R, p) and Jlog6 of its inverse (R^T, -R^T p) from one rotation log:
// log3(R^T) = -log3(R) (the axial vector changes sign exactly, the angle and
// the coefficients are the same; in the branch near pi the signs come from the
// skew part, which also flips, unless two of its entries tie exactly), so this
// equals log6_dev(R, p) and jlog6_dev(R^T, -R^T p) bit for bit.
template <class T> __device__ __forceinline__ S6<T> log6_jlog6_inv(const T *R, V3<T> p, T *Jl) {
  T t;
  V3<T> w = log3_dev(R, t);
  T al, be, db;
  log_coeffs_dev(t, al, be, db);
  jlog6_core(T(-1) * w, t, al, be, db, T(-1) * mtv(R, p), Jl);
  return {al * p - T(0.5) * cross(w, p) + (be * dot(w, p)) * w, w};
}
// The same with Jlog6(M^-1) as its blocks A, C: Jlog6 v = [A v_l + C (A v_a); A v_a]
// costs 27 products per column instead of 36, and C A is never formed.
template <class T> __device__ __forceinline__ S6<T> log6_jlog6_inv_blocks(const T *R, V3<T> p, T *A, T *C) {
  T t;
  V3<T> w = log3_dev(R, t);
  T al, be, db;
  log_coeffs_dev(t, al, be, db);
  jlog6_blocks(T(-1) * w, t, al, be, db, T(-1) * mtv(R, p), A, C);
  return {al * p - T(0.5) * cross(w, p) + (be * dot(w, p)) * w, w};
}

// y = (J J^T + damp I)^-1 e by LDL^T of the 6x6 SPD matrix, one division per
// pivot.
template <class T> __device__ __forceinline__ void ldlt6_solve(T *A, T *y) {
  T inv[6];
  for (int j = 0; j < 6; ++j) {
    T d = A[j * 6 + j];
    for (int k = 0; k < j; ++k) d -= A[j * 6 + k] * A[j * 6 + k] * A[k * 6 + k];
    A[j * 6 + j] = d;
    inv[j] = T(1) / d;
    for (int i = j + 1; i < 6; ++i) {
      T s = A[i * 6 + j];
      for (int k = 0; k < j; ++k) s -= A[i * 6 + k] * A[j * 6 + k] * A[k * 6 + k];
      A[i * 6 + j] = s * inv[j];
    }
  }
  for (int i = 0; i < 6; ++i)
    for (int k = 0; k < i; ++k) y[i] -= A[i * 6 + k] * y[k];
  for (int i = 0; i < 6; ++i) y[i] *= inv[i];
  for (int i = 5; i >= 0; --i)
    for (int k = i + 1; k < 6; ++k) y[i] -= A[k * 6 + i] * y[k];
}

// Target pose p | q (w x y z) -> R, p as Eigen::Quaterniond(w, x, y, z)
// .toRotationMatrix() builds it (pinocchio_model.cpp:255): the quaternion is
// not normalized, so a non-unit q gives (1 - |q|^2) I + |q|^2 R(q / |q|).
template <class T> __device__ __forceinline__ void pose_to_mat(const T *P, T *R, V3<T> &p) {
  const T w = P[3], x = P[4], y = P[5], z = P[6];
  R[0] = T(1) - T(2) * (y * y + z * z); R[1] = T(2) * (x * y - w * z); R[2] = T(2) * (x * z + w * y);
  R[3] = T(2) * (x * y + w * z); R[4] = T(1) - T(2) * (x * x + z * z); R[5] = T(2) * (y * z - w * x);
  R[6] = T(2) * (x * z - w * y); R[7] = T(2) * (y * z + w * x); R[8] = T(1) - T(2) * (x * x + y * y);
  p = {P[0], P[1], P[2]};
}

}  // namespace sapamd
