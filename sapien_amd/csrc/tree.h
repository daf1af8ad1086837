// Tree scans for the lanes-per-joint kernels: one environment per group of G
// lanes (G = 8, 16 or 32), lane j owning joint j of the depth-first joint
// order. Root-to-joint prefixes by pointer jumping over the 2^k-th ancestors;
// subtree sums from doubling block sums over the preorder lane ranges.
#pragma once
#include <hip/hip_runtime.h>

#include "lanes.h"
#include "spatial.h"

namespace sapamd {
namespace lanes {

template <class T> __device__ __forceinline__ S6<T> bperm6(int src4, S6<T> x) {
  return {{bperm(src4, x.l.x), bperm(src4, x.l.y), bperm(src4, x.l.z)},
          {bperm(src4, x.a.x), bperm(src4, x.a.y), bperm(src4, x.a.z)}};
}
template <class T> __device__ __forceinline__ XfT<T> bperm_xf(int src4, const XfT<T> &x) {
  XfT<T> o;
#pragma unroll
  for (int k = 0; k < 9; ++k) o.R[k] = bperm(src4, x.R[k]);
  o.p = {bperm(src4, x.p.x), bperm(src4, x.p.y), bperm(src4, x.p.z)};
  return o;
}

// log2(G) rounds of pointer jumping toward the root; T is the scalar type of
// the scanned values.
template <int G, class T = float> struct Tree {
  static constexpr int K = G == 8 ? 3 : G == 16 ? 4 : 5;
  int base4;
  int anc[K];
  int size;
  int g;
  __device__ __forceinline__ void prefix(XfT<T> &x) const {
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const int a = anc[k];
      const XfT<T> y = bperm_xf(base4 + 4 * (a < 0 ? 0 : a), x);
      x = a >= 0 ? compose(y, x) : x;
    }
  }
  __device__ __forceinline__ void prefix(S6<T> &x) const {
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const int a = anc[k];
      const S6<T> y = bperm6(base4 + 4 * (a < 0 ? 0 : a), x);
      x = sel(a >= 0, x + y, x);
    }
  }
  // value of lane g + 2^k: a DPP row shift when the group lies in one 16-lane
  // row (G <= 16; the caller masks lanes whose source is past the group),
  // ds_bpermute otherwise
  template <int k> __device__ __forceinline__ T from_lane_above(T v) const {
    if constexpr (G <= 16) {
      return from_above<(1 << k)>(v);
    } else {
      const int src = g + (1 << k);
      return bperm(base4 + 4 * (src < G ? src : 0), v);
    }
  }
  // subtree sum of x over [g, g + size) (x = 0 on lanes without a joint)
  template <int N> __device__ __forceinline__ void subtree(T (&x)[N]) const {
    T blk[K][N];
    T cur[N];
#pragma unroll
    for (int i = 0; i < N; ++i) cur[i] = x[i];
#pragma unroll
    for (int k = 0; k < K; ++k) {
#pragma unroll
      for (int i = 0; i < N; ++i) blk[k][i] = cur[i];
      const bool in = g + (1 << k) < G;
#pragma unroll
      for (int i = 0; i < N; ++i) {
        T y;
        switch (k) {  // k is a constant once the loop is unrolled
          case 0: y = from_lane_above<0>(cur[i]); break;
          case 1: y = from_lane_above<1>(cur[i]); break;
          case 2: y = from_lane_above<2>(cur[i]); break;
          case 3: y = from_lane_above<3>(cur[i]); break;
          default: y = from_lane_above<4>(cur[i]); break;
        }
        cur[i] = in ? cur[i] + y : cur[i];
      }
    }
    // blocks of size 2^K cover the whole group: size <= G
    T acc[N];
#pragma unroll
    for (int i = 0; i < N; ++i) acc[i] = T(0);
    // every lane runs the query (ds_bpermute must not sit in divergent code);
    // size == G (a single-root tree filling the group) has no bits below K and
    // takes the full block instead
    int pos = g;
#pragma unroll
    for (int k = K - 1; k >= 0; --k) {
      const bool take = (size >> k) & 1;
      const int s4 = base4 + 4 * (pos < G ? pos : 0);
#pragma unroll
      for (int i = 0; i < N; ++i) {
        const T y = bperm(s4, blk[k][i]);
        acc[i] = take ? acc[i] + y : acc[i];
      }
      pos += take ? (1 << k) : 0;
    }
#pragma unroll
    for (int i = 0; i < N; ++i) x[i] = size >= G ? cur[i] : acc[i];
  }
  __device__ __forceinline__ void subtree(S6<T> &f) const {
    T x[6] = {f.l.x, f.l.y, f.l.z, f.a.x, f.a.y, f.a.z};
    subtree<6>(x);
    f = {{x[0], x[1], x[2]}, {x[3], x[4], x[5]}};
  }
};

// Sum over the G lanes of the group, result in lane 0.
template <int G> __device__ __forceinline__ float group_sum(int g, int base4, float x) {
#pragma unroll
  for (int d = 1; d < G; d *= 2) {
    const int src = g + d;
    const float y = bperm(base4 + 4 * (src < G ? src : 0), x);
    x = src < G ? x + y : x;
  }
  return x;
}

}  // namespace lanes
}  // namespace sapamd
