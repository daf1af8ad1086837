// Lanes-per-joint CLIK: damped least squares on the SE(3) log, as
// compute_inverse_kinematics (pinocchio_model.cpp:228-300), for problems whose
// root-to-link joint chain has <= 16 joints. One problem per group of G = 8 or
// 16 lanes inside a 16-lane row; lane k owns chain joint k.
//   FK      each lane's local transform, prefix product by DPP row shifts; the
//           link joint's pose broadcast to the group
//   error   log6 of oMj^-1 oMdes (every lane, same bits)
//   J J^T   each lane's Jacobian column in the link frame through Jlog6; the 21
//           products in FP64 (exact for FP32 factors) summed by a butterfly
//           over the group, so every lane holds the same sum
//   solve   the damped 6x6 LDL^T in FP64, redundantly in every lane
// Wavefronts pull problems from a work queue (one index per group) and each
// group leaves the iteration loop on its own problem's exit test: per-problem
// early exit. Results follow the
// thread-per-problem kernel (kernels_generic.hip) step for step; only the
// association order of the FK products and of the J J^T sums differs.
#include <hip/hip_runtime.h>

#include "lanes.h"
#include "launch.h"
#include "se3.h"
#include "spatial.h"
#include "tree.h"

namespace sapamd {

namespace {

using namespace lanes;

// 32-bit halves of a double moved by one DPP pattern (a full permutation of the
// 16-lane row, so every lane has a source).
template <int CTRL> __device__ __forceinline__ double dpp_d(double x) {
  const uint64_t u = __builtin_bit_cast(uint64_t, x);
  const uint32_t lo = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)u, CTRL, 0xF, 0xF, false);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)(u >> 32), CTRL, 0xF, 0xF, false);
  return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}

// Sum over the G lanes of the group, in every lane. Each round adds the value
// of a partner lane (quad_perm xor 1, xor 2, half-row mirror, row mirror);
// partners compute a + b and b + a, which are equal, so all lanes end with the
// same bits.
template <int G> __device__ __forceinline__ double group_allsum(double x) {
  x += dpp_d<0xB1>(x);   // quad_perm [1,0,3,2]
  x += dpp_d<0x4E>(x);   // quad_perm [2,3,0,1]
  x += dpp_d<0x141>(x);  // row_half_mirror: i <-> 7 - i
  if constexpr (G == 16) x += dpp_d<0x140>(x);  // row_mirror: i <-> 15 - i
  return x;
}

__device__ __forceinline__ int bperm_i(int src4, int x) { return __builtin_amdgcn_ds_bpermute(src4, x); }

// One problem of segment S on this group (all G lanes active).
template <int G>
__device__ void clik_group(const ClikSegment &S, int64_t prob, int g, int base4) {
  const ClikParams &P = S.P;
  const ModelTable<float> *M = S.tab;
  const int nv = M->nv, nc = P.nchain;
  const int64_t tgt = P.ntargets ? prob % P.ntargets : prob;
  const int32_t start = P.ntargets ? (int32_t)(prob / P.ntargets) : 0;
  float *qo = S.q_out + prob * nv;
  const float *q0 = S.q0 ? S.q0 + prob * nv : nullptr;
  if (P.solved) {
    // a lower start already converged for this target: skip (lane 0's reading
    // decides for the group; every lane loads, so the permute reads no value
    // produced under a divergent branch)
    int lowest = __hip_atomic_load(P.solved + tgt, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    lowest = bperm_i(base4, lowest);
    if (lowest < start) {
      for (int u = g; u < nv; u += G) qo[u] = q0 ? q0[u] : 0.f;
      if (g == 0) {
        S.success[prob] = 0;
        for (int r = 0; r < 6; ++r) S.err[prob * 6 + r] = 1e30f;
        if (S.iters) S.iters[prob] = -1;
      }
      return;
    }
  }
  const bool act = g < nc;
  const JointRec<float> &Jg = joints_of(M)[P.chain[act ? g : 0]];
  const int kind = Jg.kind, vu = Jg.vuser;
  float JA[9], JB[9], JC[9];
#pragma unroll
  for (int k = 0; k < 9; ++k) {
    JA[k] = Jg.A[k];
    JB[k] = Jg.Bq[k];
    JC[k] = Jg.Cq[k];
  }
  const V3<float> Jp = load3(Jg.p), Jap = load3(Jg.ap), Jax = load3(Jg.axis);

  // desired pose of the link's joint frame: oMdes = target * link2joint^-1
  const LinkRec<float> &Lk = links_of(M)[P.link];
  float Rt[9], Rd[9];
  V3<float> pt;
  pose_to_mat(S.target + tgt * 7, Rt, pt);
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j)
      Rd[3 * i + j] = Rt[3 * i] * Lk.R[3 * j] + Rt[3 * i + 1] * Lk.R[3 * j + 1] + Rt[3 * i + 2] * Lk.R[3 * j + 2];
  const V3<float> pd = pt - mv(Rd, load3(Lk.p));

  float qc = act ? (q0 ? q0[vu] : 0.f) : 0.f, qbest = qc;
  const float msk = act ? (P.mask ? float(P.mask[prob * P.mask_ld + vu]) : 1.f) : 0.f;
  float minErr = 1e10f, bestErr[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
  bool ok = false;
  int it = 0;
  const int last4 = base4 + 4 * (nc > 0 ? nc - 1 : 0);
  for (;; ++it) {
    // FK of the chain: local transform, then the prefix product over the group
    // (selects, not branches: no divergence inside the group before the scan)
    Xf X;
    {
      float s, c;
      sincosf(qc, &s, &c);
      const bool pri = kind == KIND_PRISMATIC;
      const float sr = pri ? 0.f : s, omc = pri ? 0.f : 1.f - c, qp = pri ? qc : 0.f;
#pragma unroll
      for (int k = 0; k < 9; ++k) X.R[k] = JA[k] + sr * JB[k] + omc * JC[k];
      X.p = Jp + qp * Jap;
      if (!act) X = identity_xf();
    }
    scan_xf<G>(g, X);
    Xf E = bperm_xf(last4, X);  // the link joint's world pose
    if (nc == 0) E = identity_xf();

    // iMd = oMj^-1 oMdes and its log
    float Ri[9];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) Ri[3 * i + j] = E.R[i] * Rd[j] + E.R[3 + i] * Rd[3 + j] + E.R[6 + i] * Rd[6 + j];
    const V3<float> pi = mtv(E.R, pd - E.p);
    const S6<float> e = log6_dev(Ri, pi);
    float ev[6];
    vec_of(e, ev);
    const float en = sqrtf(dot6(e, e));
    if (en < minErr) {
      minErr = en;
      qbest = qc;
#pragma unroll
      for (int r = 0; r < 6; ++r) bestErr[r] = ev[r];
    }
    if (en < float(P.eps)) {
      ok = true;
      break;
    }
    if (it >= P.max_iter) break;

    // this lane's column of Jlog6(iMd^-1) * J_local, masked
    float Rinv[9];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) Rinv[3 * i + j] = Ri[3 * j + i];
    float Jl[36];
    jlog6_dev(Rinv, -1.f * mtv(Ri, pi), Jl);
    S6<float> cw = zero6<float>();
    if (act) cw = kind == KIND_PRISMATIC ? S6<float>{mv(X.R, Jax), {0.f, 0.f, 0.f}}
                                         : S6<float>{cross(X.p, mv(X.R, Jax)), mv(X.R, Jax)};
    float cv[6];
    vec_of(to_local(E.R, E.p, cw), cv);
    float Jk[6];
#pragma unroll
    for (int r = 0; r < 6; ++r) {
      float acc = 0.f;
#pragma unroll
      for (int m = 0; m < 6; ++m) acc += Jl[r * 6 + m] * cv[m];
      Jk[r] = -acc * msk;
    }
    // J J^T + damp I and its solve in FP64 (cond ~1e6+ near singular poses)
    double JJt[36], y[6];
#pragma unroll
    for (int r = 0; r < 6; ++r)
#pragma unroll
      for (int m = 0; m <= r; ++m) JJt[r * 6 + m] = group_allsum<G>(double(Jk[r]) * double(Jk[m]));
#pragma unroll
    for (int r = 0; r < 6; ++r) {
      JJt[r * 6 + r] += P.damp;
#pragma unroll
      for (int m = 0; m < r; ++m) JJt[m * 6 + r] = JJt[r * 6 + m];
      y[r] = double(ev[r]);
    }
    ldlt6_solve(JJt, y);
    float v = 0.f;
#pragma unroll
    for (int r = 0; r < 6; ++r) v -= Jk[r] * float(y[r]);
    qc += v * float(P.dt);
  }

  // result: chain dofs from the best iterate (continuous wrapped), the rest from q0
  for (int u = g; u < nv; u += G) {
    bool chain_dof = false;
    for (int k = 0; k < nc; ++k) chain_dof |= joints_of(M)[P.chain[k]].vuser == u;
    if (!chain_dof) qo[u] = q0 ? q0[u] : 0.f;
  }
  if (act) {
    float v = qbest;
    if (kind == KIND_CONTINUOUS) {
      float s, c;
      sincosf(v, &s, &c);
      v = atan2f(s, c);
    }
    qo[vu] = v;
  }
  if (g == 0) {
    S.success[prob] = ok ? 1 : 0;
#pragma unroll
    for (int r = 0; r < 6; ++r) S.err[prob * 6 + r] = bestErr[r];
    if (S.iters) S.iters[prob] = it;
    if (ok && P.solved) __hip_atomic_fetch_min(P.solved + tgt, start, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}

// Persistent: each wavefront takes the next 64 / G problem indices from the
// queue (one per group) until it runs dry; the segment table sits in LDS
// (indexed per group). The fetch is wavefront-wide on purpose: the groups of a
// wavefront share one instruction stream, so a group cannot start its next
// problem before the others finish theirs anyway, and every lane executes the
// atomic, so the fetched index never travels from a lane that took a
// different branch.
template <int G> __global__ __launch_bounds__(256) void k_clik_lanes(HeteroClikArgs A) {
  __shared__ __align__(16) int s_raw[sizeof(ClikSegment) * kMaxSegments / 4];
  {
    const int *src = reinterpret_cast<const int *>(A.seg);
    for (int i = threadIdx.x; i < (int)(sizeof(s_raw) / 4); i += blockDim.x) s_raw[i] = src[i];
  }
  __syncthreads();
  const ClikSegment *s_seg = reinterpret_cast<const ClikSegment *>(s_raw);
  const int lane = threadIdx.x & 63;
  const int g = lane & (G - 1);
  const int base4 = (lane & ~(G - 1)) * 4;
  constexpr int kPerWave = 64 / G;
  for (;;) {
    // all 64 lanes active here: lane 0 adds kPerWave, the others add 0, so lane
    // 0's old value is this wavefront's first index
    const int old = atomicAdd(A.counter, lane == 0 ? kPerWave : 0);
    const int first = __builtin_amdgcn_readlane(old, 0);
    if ((int64_t)first >= A.total) break;  // uniform over the wavefront
    const int idx = first + lane / G;
    if ((int64_t)idx < A.total) {
      int s = 0;
      while (s + 1 < A.nseg && idx >= s_seg[s + 1].first) ++s;
      clik_group<G>(s_seg[s], idx - s_seg[s].first, g, base4);
    }
  }
}

}  // namespace

int clik_lanes_group(int max_chain) { return max_chain <= 8 ? 8 : max_chain <= 16 ? 16 : 0; }

hipError_t launch_clik_lanes(const HeteroClikArgs &A, int max_chain, int blocks, hipStream_t s) {
  if (A.total <= 0) return hipSuccess;
  const int G = clik_lanes_group(max_chain);
  if (!G || A.total > INT32_MAX) return hipErrorNotSupported;
  if (G == 8) hipLaunchKernelGGL(k_clik_lanes<8>, dim3((unsigned)blocks), dim3(256), 0, s, A);
  else hipLaunchKernelGGL(k_clik_lanes<16>, dim3((unsigned)blocks), dim3(256), 0, s, A);
  return hipGetLastError();
}

}  // namespace sapamd
