// Cross-lane building blocks for the lanes-per-joint kernels: DPP row shifts,
// ds_bpermute, SE(3) / spatial-vector scans over a group of G lanes inside a
// 16-lane row, and per-joint helpers shared by the forward and adjoint kernels.
#pragma once
#include <hip/hip_runtime.h>

#include "spatial.h"

namespace sapamd {
namespace lanes {

template <int CTRL> __device__ __forceinline__ float dppf(float x) {
  return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, x), CTRL, 0xF, 0xF, true));
}
// value held by lane g - D of the same 16-lane row (0 for lanes below D)
template <int D> __device__ __forceinline__ float from_below(float x) { return dppf<0x110 + D>(x); }
// value held by lane g + D of the same 16-lane row (0 past the row end)
template <int D> __device__ __forceinline__ float from_above(float x) { return dppf<0x100 + D>(x); }
template <int D> __device__ __forceinline__ double from_above(double x) {
  const uint64_t u = __builtin_bit_cast(uint64_t, x);
  const uint32_t lo = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)u, 0x100 + D, 0xF, 0xF, true);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)(u >> 32), 0x100 + D, 0xF, 0xF, true);
  return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}
// value held by wave lane src4 / 4
__device__ __forceinline__ float bperm(int src4, float x) {
  return __builtin_bit_cast(float, __builtin_amdgcn_ds_bpermute(src4, __builtin_bit_cast(int, x)));
}
__device__ __forceinline__ double bperm(int src4, double x) {
  const uint64_t u = __builtin_bit_cast(uint64_t, x);
  const uint32_t lo = (uint32_t)__builtin_amdgcn_ds_bpermute(src4, (int)(uint32_t)u);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_ds_bpermute(src4, (int)(uint32_t)(u >> 32));
  return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}

template <int D> __device__ __forceinline__ S6<float> below6(S6<float> v) {
  return {{from_below<D>(v.l.x), from_below<D>(v.l.y), from_below<D>(v.l.z)},
          {from_below<D>(v.a.x), from_below<D>(v.a.y), from_below<D>(v.a.z)}};
}
template <int D> __device__ __forceinline__ S6<float> above6(S6<float> v) {
  return {{from_above<D>(v.l.x), from_above<D>(v.l.y), from_above<D>(v.l.z)},
          {from_above<D>(v.a.x), from_above<D>(v.a.y), from_above<D>(v.a.z)}};
}
template <class T> __device__ __forceinline__ S6<T> sel(bool t, S6<T> a, S6<T> b) {
  return {{t ? a.l.x : b.l.x, t ? a.l.y : b.l.y, t ? a.l.z : b.l.z},
          {t ? a.a.x : b.a.x, t ? a.a.y : b.a.y, t ? a.a.z : b.a.z}};
}

// Orders one wavefront's LDS accesses across lanes: the fences keep the
// compiler from moving loads or stores across the point (a wavefront's own LDS
// operations complete in order, so no wait is emitted at wavefront scope).
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// sin and cos of a joint angle. For |x| < 16 rad the hardware sine / cosine
// (v_sin_f32 / v_cos_f32, argument scaled to revolutions): absolute error
// below 1e-6 there, measured against the FP64 oracle through the kernels'
// outputs. Larger angles (continuous joints far from zero) take the
// full-range library routine.
__device__ __forceinline__ void joint_sincos(float x, float *s, float *c) {
  if (fabsf(x) < 16.f) {
    *s = __sinf(x);
    *c = __cosf(x);
  } else {
    sincosf(x, s, c);
  }
}
__device__ __forceinline__ void joint_sincos(double x, double *s, double *c) { sincos(x, s, c); }

template <class T> struct XfT {
  T R[9];
  V3<T> p;
};
using Xf = XfT<float>;
template <class T> __device__ __forceinline__ XfT<T> compose(const XfT<T> &a, const XfT<T> &b) {
  XfT<T> o;
  mm(a.R, b.R, o.R);
  o.p = mv(a.R, b.p) + a.p;
  return o;
}
template <class T = float> __device__ __forceinline__ XfT<T> identity_xf() {
  XfT<T> o;
#pragma unroll
  for (int k = 0; k < 9; ++k) o.R[k] = (k % 4 == 0) ? T(1) : T(0);
  o.p = {T(0), T(0), T(0)};
  return o;
}
template <int D> __device__ __forceinline__ Xf below_xf(const Xf &x) {
  Xf o;
#pragma unroll
  for (int k = 0; k < 9; ++k) o.R[k] = from_below<D>(x.R[k]);
  o.p = {from_below<D>(x.p.x), from_below<D>(x.p.y), from_below<D>(x.p.z)};
  return o;
}
__device__ __forceinline__ Xf sel_xf(bool t, const Xf &a, const Xf &b) {
  Xf o;
#pragma unroll
  for (int k = 0; k < 9; ++k) o.R[k] = t ? a.R[k] : b.R[k];
  o.p = {t ? a.p.x : b.p.x, t ? a.p.y : b.p.y, t ? a.p.z : b.p.z};
  return o;
}

// Inclusive prefix product over the group: x_g <- x_0 ... x_g.
template <int G, int D = 1> __device__ __forceinline__ void scan_xf(int g, Xf &x) {
  if constexpr (D < G) {
    Xf y = below_xf<D>(x);
    x = sel_xf(g >= D, compose(y, x), x);
    scan_xf<G, 2 * D>(g, x);
  }
}
template <int G, int D = 1> __device__ __forceinline__ void scan_sum(int g, S6<float> &x) {
  if constexpr (D < G) {
    x = sel(g >= D, x + below6<D>(x), x);
    scan_sum<G, 2 * D>(g, x);
  }
}
template <int G, int D = 1> __device__ __forceinline__ void scan_suffix(int g, S6<float> &x) {
  if constexpr (D < G) {
    x = sel(g + D < G, x + above6<D>(x), x);
    scan_suffix<G, 2 * D>(g, x);
  }
}
// exclusive values from the inclusive scan
template <int G> __device__ __forceinline__ Xf excl_xf(int g, const Xf &incl) {
  if constexpr (G == 1) return identity_xf();
  else return sel_xf(g >= 1, below_xf<1>(incl), identity_xf());
}
template <int G> __device__ __forceinline__ S6<float> excl_sum(int g, S6<float> incl) {
  if constexpr (G == 1) return zero6<float>();
  else return sel(g >= 1, below6<1>(incl), zero6<float>());
}
template <int G> __device__ __forceinline__ S6<float> excl_suffix(int g, S6<float> incl) {
  if constexpr (G == 1) return zero6<float>();
  else return sel(g + 1 < G, above6<1>(incl), zero6<float>());
}

// Interleaved pairs (8 lanes per environment): the two environments of a
// 16-lane row take the even and the odd lanes, joint j of environment p on
// lane 2j + p. A step of D joints is then a DPP row shift of 2D lanes: it never
// mixes the two environments, and a source past the row edge yields the
// instruction's `old` operand (identity entry or zero), so the scans need no
// selects.
template <int CTRL> __device__ __forceinline__ float dpp_old(float old, float x) {
  return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(__builtin_bit_cast(int, old), __builtin_bit_cast(int, x),
                                                               CTRL, 0xF, 0xF, false));
}
// (The zero entries could use bound_ctrl DPP moves, which need no `old`
// value first: 27 fewer instructions per wavefront, but measured twice as slow
// at 262144 environments and more on gfx950, so the explicit `old` stays.)
template <int D> __device__ __forceinline__ Xf below_xf_il(const Xf &x) {  // joint j - D, identity before joint 0
  Xf o;
#pragma unroll
  for (int k = 0; k < 9; ++k) o.R[k] = dpp_old<0x110 + 2 * D>((k % 4 == 0) ? 1.f : 0.f, x.R[k]);
  o.p = {dpp_old<0x110 + 2 * D>(0.f, x.p.x), dpp_old<0x110 + 2 * D>(0.f, x.p.y), dpp_old<0x110 + 2 * D>(0.f, x.p.z)};
  return o;
}
template <int D> __device__ __forceinline__ S6<float> below6_il(S6<float> v) {
  return {{dpp_old<0x110 + 2 * D>(0.f, v.l.x), dpp_old<0x110 + 2 * D>(0.f, v.l.y), dpp_old<0x110 + 2 * D>(0.f, v.l.z)},
          {dpp_old<0x110 + 2 * D>(0.f, v.a.x), dpp_old<0x110 + 2 * D>(0.f, v.a.y), dpp_old<0x110 + 2 * D>(0.f, v.a.z)}};
}
template <int D> __device__ __forceinline__ S6<float> above6_il(S6<float> v) {
  return {{dpp_old<0x100 + 2 * D>(0.f, v.l.x), dpp_old<0x100 + 2 * D>(0.f, v.l.y), dpp_old<0x100 + 2 * D>(0.f, v.l.z)},
          {dpp_old<0x100 + 2 * D>(0.f, v.a.x), dpp_old<0x100 + 2 * D>(0.f, v.a.y), dpp_old<0x100 + 2 * D>(0.f, v.a.z)}};
}
// inclusive prefix product / prefix sum / suffix sum over the 8 joints
template <int D = 1> __device__ __forceinline__ void scan_xf_il(Xf &x) {
  if constexpr (D < 8) {
    x = compose(below_xf_il<D>(x), x);
    scan_xf_il<2 * D>(x);
  }
}
template <int D = 1> __device__ __forceinline__ void scan_sum_il(S6<float> &x) {
  if constexpr (D < 8) {
    x = x + below6_il<D>(x);
    scan_sum_il<2 * D>(x);
  }
}
template <int D = 1> __device__ __forceinline__ void scan_suffix_il(S6<float> &x) {
  if constexpr (D < 8) {
    x = x + above6_il<D>(x);
    scan_suffix_il<2 * D>(x);
  }
}
template <int D = 1> __device__ __forceinline__ void scan_suffix1_il(float &x) {  // total on joint 0's lane
  if constexpr (D < 8) {
    x = x + dpp_old<0x100 + 2 * D>(0.f, x);
    scan_suffix1_il<2 * D>(x);
  }
}

// Shoemake's quaternion with Eigen's branch rule (w first when the trace is
// positive, else the largest diagonal entry), evaluated without divergence.
__device__ __forceinline__ void quat_branchless(const float *R, float *q) {
  float tr = R[0] + R[4] + R[8];
  int i = R[4] > R[0] ? 1 : 0;
  float di = i ? R[4] : R[0];
  int k = tr > 0.f ? 0 : 1 + (R[8] > di ? 2 : i);
  float t = k == 0 ? tr : k == 1 ? R[0] - R[4] - R[8] : k == 2 ? R[4] - R[8] - R[0] : R[8] - R[0] - R[4];
  // t + 1 >= 1 on the chosen branch: the hardware reciprocal square root
  // (about 1 ulp) replaces the correctly rounded sqrt and division
  const float rs = __builtin_amdgcn_rsqf(t + 1.f);
  float r = (t + 1.f) * rs;
  float h = 0.5f * rs;
  float a = (R[7] - R[5]) * h, b = (R[2] - R[6]) * h, c = (R[3] - R[1]) * h;  // w-type
  float d = (R[1] + R[3]) * h, e = (R[2] + R[6]) * h, f = (R[5] + R[7]) * h;  // off-diagonal sums
  float big = 0.5f * r;
  q[0] = k == 0 ? big : k == 1 ? a : k == 2 ? b : c;
  q[1] = k == 0 ? a : k == 1 ? big : k == 2 ? d : e;
  q[2] = k == 0 ? b : k == 1 ? d : k == 2 ? big : f;
  q[3] = k == 0 ? c : k == 1 ? e : k == 2 ? f : big;
}

// Output store; NT = non-temporal (streamed past the caches). With 8 or more
// lanes per environment the lanes of a wave cover whole runs of each output
// row, and streaming them is faster; with fewer lanes the runs are scattered
// and cached stores merge better.
template <bool NT> __device__ __forceinline__ void st(float v, float *p) {
  if constexpr (NT) __builtin_nontemporal_store(v, p);
  else *p = v;
}

// The wavefront's lanes copy n floats from its LDS run to global memory: with
// a 16-byte aligned destination (vec4; the LDS runs always are) 64 float4 per
// store instruction, else 64 consecutive floats per instruction.
template <bool NT> __device__ __forceinline__ void copy_run(const float *src, float *dst, int n, int lane, bool vec4) {
  int i0 = 0;
  if (vec4) {
    const int n4 = n >> 2;
    typedef float v4f __attribute__((ext_vector_type(4)));
    const v4f *s4 = reinterpret_cast<const v4f *>(src);
    v4f *d4 = reinterpret_cast<v4f *>(dst);
    for (int i = lane; i < n4; i += 64) {
      if constexpr (NT) __builtin_nontemporal_store(s4[i], d4 + i);
      else d4[i] = s4[i];
    }
    i0 = n4 << 2;
  }
  for (int i = i0 + lane; i < n; i += 64) st<NT>(src[i], dst + i);
}

// A constant pose (position, quaternion already computed).
template <bool NT = false> __device__ __forceinline__ void store_pose_q(float *P, const float *p, const float *q) {
  st<NT>(p[0], P + 0); st<NT>(p[1], P + 1); st<NT>(p[2], P + 2);
  st<NT>(q[0], P + 3); st<NT>(q[1], P + 4); st<NT>(q[2], P + 5); st<NT>(q[3], P + 6);
}

template <bool NT = false> __device__ __forceinline__ void store_pose(float *P, const float *R, V3<float> p) {
  float qq[4];
  quat_branchless(R, qq);
  st<NT>(p.x, P + 0); st<NT>(p.y, P + 1); st<NT>(p.z, P + 2);
  st<NT>(qq[0], P + 3); st<NT>(qq[1], P + 4); st<NT>(qq[2], P + 5); st<NT>(qq[3], P + 6);
}

// World inertia of a body whose rotational inertia is diagonal in its frame:
// R diag(d) R^T = sum_k d_k r_k r_k^T (columns r_k of R).
template <class T>
__device__ __forceinline__ WInertia<T> world_inertia_fast(const JointRec<T> &J, const T *R, V3<T> p) {
  if (!J.diagI) return world_inertia(J, R, p);
  WInertia<T> Y;
  Y.m = J.m;
  Y.c = mv(R, load3(J.c)) + p;
  const T d0 = J.Ic[0], d1 = J.Ic[3], d2 = J.Ic[5];
  Y.I[0] = d0 * R[0] * R[0] + d1 * R[1] * R[1] + d2 * R[2] * R[2];
  Y.I[1] = d0 * R[0] * R[3] + d1 * R[1] * R[4] + d2 * R[2] * R[5];
  Y.I[2] = d0 * R[0] * R[6] + d1 * R[1] * R[7] + d2 * R[2] * R[8];
  Y.I[3] = d0 * R[3] * R[3] + d1 * R[4] * R[4] + d2 * R[5] * R[5];
  Y.I[4] = d0 * R[3] * R[6] + d1 * R[4] * R[7] + d2 * R[5] * R[8];
  Y.I[5] = d0 * R[6] * R[6] + d1 * R[7] * R[7] + d2 * R[8] * R[8];
  return Y;
}


// Scalar sums over the group (the total lands in lane 0 of the group).
template <int G, int D = 1> __device__ __forceinline__ void scan_suffix1(int g, float &x) {
  if constexpr (D < G) {
    const float y = from_above<D>(x);
    x = g + D < G ? x + y : x;
    scan_suffix1<G, 2 * D>(g, x);
  }
}

}  // namespace lanes
}  // namespace sapamd
