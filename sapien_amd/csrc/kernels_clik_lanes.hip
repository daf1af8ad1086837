// Lanes-per-joint CLIK: damped least squares on the SE(3) log, as
// compute_inverse_kinematics (pinocchio_model.cpp:228-300), for problems whose
// root-to-link joint chain has <= 16 joints. One problem per group of G lanes
// inside a 16-lane row, C consecutive chain joints per lane (G x C >= chain).
//   FK      each lane composes its C local transforms; a DPP prefix product
//           over the group gives the world transform before the lane's first
//           joint; the link joint's pose (the inclusive product at its lane:
//           joints past the chain end are identities) goes to the whole group
//   error   log6 of oMj^-1 oMdes (every lane, same bits)
//   J J^T   each joint's column in the link frame through Jlog6; the 21
//           products in FP64 (exact for FP32 factors), summed over the lane's
//           joints and then by a butterfly over the group (every lane ends
//           with the same bits)
//   solve   the damped 6x6 LDL^T in FP64, redundantly in every lane
// Scheduling: one trip of the main loop is one CLIK iteration for every group
// of the wavefront. A group whose problem ends (converged or out of
// iterations) writes its result and takes the next problem at the following
// trip, so one slow problem does not hold its neighbours idle. The queue is
// fetched once per trip for all groups that need work (ballot, one atomic
// from every lane), and every cross-lane operation sits in code all lanes
// execute.
#include <hip/hip_runtime.h>

#include "lanes.h"
#include "launch.h"
#include "se3.h"
#include "spatial.h"
#include "tree.h"

namespace sapamd {

namespace {

using namespace lanes;

constexpr int kClikMaxChain = 16;

// The chain joints' constants the loop reads, staged in LDS per segment. The
// staged form is branch-free: Bq = Cq = 0 for prismatic joints, ap = 0 for the
// others, and slot kClikMaxChain is an identity record (axis 0) that every
// lane past the chain end reads.
struct ChainRec {
  float A[9], Bq[9], Cq[9];
  float p[3], ap[3], axis[3];
  int32_t kind, vuser, pad[2];
};

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
  if constexpr (G >= 2) x += dpp_d<0xB1>(x);    // quad_perm [1,0,3,2]
  if constexpr (G >= 4) x += dpp_d<0x4E>(x);    // quad_perm [2,3,0,1]
  if constexpr (G >= 8) x += dpp_d<0x141>(x);   // row_half_mirror: i <-> 7 - i
  if constexpr (G >= 16) x += dpp_d<0x140>(x);  // row_mirror: i <-> 15 - i
  return x;
}

__device__ __forceinline__ int bperm_i(int src4, int x) { return __builtin_amdgcn_ds_bpermute(src4, x); }

// Local transform of a chain joint from its staged record (no branches and no
// selects: the record's zero entries drop the other joint kind's terms, and
// the identity record gives the identity for q = 0).
__device__ __forceinline__ Xf local_xf(const ChainRec &J, float q) {
  float s, c;
  joint_sincos(q, &s, &c);
  const float omc = 1.f - c;
  Xf X;
#pragma unroll
  for (int k = 0; k < 9; ++k) X.R[k] = J.A[k] + s * J.Bq[k] + omc * J.Cq[k];
  X.p = load3(J.p) + q * load3(J.ap);
  return X;
}

// Per-group problem state (the same in every lane of the group, except the
// per-joint entries).
template <int C> struct ClikState {
  int64_t prob;  // problem index within its segment
  int seg, nc, it;
  float Rd[9];
  V3<float> pd;
  float q[C], qbest[C], msk[C];
  float minErr, bestErr[6];
  float midErr;  // error at iteration cap / 2 (first phase)
};

enum : int { kIdle = 0, kRunning = 1, kRetired = 2 };

// CONT = false: the first phase (problems from A.counter in index order; those
// still running after A.cap iterations are handed to A.cont). CONT = true: the
// second phase (the handed-over problems, from A.cont_queue).
template <int G, int C, bool CONT> __global__ __launch_bounds__(256) void k_clik_lanes(HeteroClikArgs A) {
  __shared__ __align__(16) int s_raw[sizeof(ClikSegment) * kMaxSegments / 4];
  __shared__ ChainRec s_chain[kMaxSegments][kClikMaxChain + 1];
  {
    const int *src = reinterpret_cast<const int *>(A.seg);
    for (int i = threadIdx.x; i < (int)(sizeof(s_raw) / 4); i += blockDim.x) s_raw[i] = src[i];
    constexpr int kSlots = kClikMaxChain + 1;
    for (int i = threadIdx.x; i < A.nseg * kSlots; i += blockDim.x) {
      const ClikSegment &S = A.seg[i / kSlots];
      const int k = i % kSlots;
      ChainRec &R = s_chain[i / kSlots][k];
      if (k == kClikMaxChain) {  // identity record
        for (int t = 0; t < 9; ++t) {
          R.A[t] = (t % 4 == 0) ? 1.f : 0.f;
          R.Bq[t] = R.Cq[t] = 0.f;
        }
        for (int t = 0; t < 3; ++t) R.p[t] = R.ap[t] = R.axis[t] = 0.f;
        R.kind = KIND_REVOLUTE;
        R.vuser = 0;
        continue;
      }
      if (k >= S.P.nchain) continue;
      const JointRec<float> &J = joints_of(S.tab)[S.P.chain[k]];
      const bool pri = J.kind == KIND_PRISMATIC;
      for (int t = 0; t < 9; ++t) {
        R.A[t] = J.A[t];
        R.Bq[t] = pri ? 0.f : J.Bq[t];
        R.Cq[t] = pri ? 0.f : J.Cq[t];
      }
      for (int t = 0; t < 3; ++t) {
        R.p[t] = J.p[t];
        R.ap[t] = pri ? J.ap[t] : 0.f;
        R.axis[t] = J.axis[t];
      }
      R.kind = J.kind;
      R.vuser = J.vuser;
    }
  }
  __syncthreads();
  const ClikSegment *s_seg = reinterpret_cast<const ClikSegment *>(s_raw);

  const int lane = threadIdx.x & 63;
  const int g = lane & (G - 1);
  const int glane0 = lane & ~(G - 1);  // first lane of the group
  const int base4 = glane0 * 4;
  const uint64_t below = glane0 ? ((uint64_t)1 << glane0) - 1 : 0;

  ClikState<C> st;
  st.prob = 0;
  st.seg = 0;
  st.nc = 0;
  st.it = 0;
#pragma unroll
  for (int k = 0; k < 9; ++k) st.Rd[k] = (k % 4 == 0) ? 1.f : 0.f;
  st.pd = {0.f, 0.f, 0.f};
#pragma unroll
  for (int c = 0; c < C; ++c) st.q[c] = st.qbest[c] = st.msk[c] = 0.f;
  st.minErr = 1e10f;
  st.midErr = 0.f;
#pragma unroll
  for (int r = 0; r < 6; ++r) st.bestErr[r] = 0.f;
  int state = kIdle;
  // problems in this phase's queue (the second phase reads the first's counts)
  const int64_t nhard =
      CONT ? (int64_t)__hip_atomic_load(A.cont_count, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : 0;
  const int64_t total =
      CONT ? nhard + __hip_atomic_load(A.cont_easy, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : A.total;

  for (;;) {
    // ---- refill: the groups without a problem take the next indices
    const uint64_t want = __ballot(state == kIdle && g == 0);
    if (want) {  // uniform
      const int nwant = __popcll(want);
      const int old = atomicAdd(CONT ? A.cont_queue : A.counter, lane == 0 ? nwant : 0);
      const int first = __builtin_amdgcn_readlane(old, 0);
      const int64_t idx = (int64_t)first + __popcll(want & below);
      // second phase: the front records first, then the back ones
      const ClikCont *rec =
          CONT ? A.cont + (idx < nhard ? idx : idx < total ? A.total - 1 - (idx - nhard) : 0) : nullptr;
      // multi-start skip test (first phase): every lane loads (from a valid
      // address in any case) and the group takes its first lane's reading
      const int64_t ic = CONT ? (idx < total ? rec->prob : 0) : (idx < total ? idx : 0);
      int s = 0;
      while (s + 1 < A.nseg && ic >= s_seg[s + 1].first) ++s;
      const ClikSegment &S = s_seg[s];
      const ClikParams &P = S.P;
      const int64_t prob = ic - S.first;
      const int64_t tgt = P.ntargets ? prob % P.ntargets : prob;
      const int32_t start = P.ntargets ? (int32_t)(prob / P.ntargets) : 0;
      int lowest = INT32_MAX;
      if constexpr (!CONT) {
        lowest = __hip_atomic_load(P.solved ? P.solved + tgt : A.counter, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        lowest = bperm_i(base4, lowest);
      }
      if (state == kIdle) {  // per group; no cross-lane operations inside
        if (idx >= total) {
          state = kRetired;
        } else if (!CONT && P.solved && lowest < start) {
          // a lower start already converged for this target: skip
          const int nv = S.tab->nv;
          float *qo = S.q_out + prob * nv;
          const float *q0 = S.q0 ? S.q0 + prob * nv : nullptr;
          for (int u = g; u < nv; u += G) qo[u] = q0 ? q0[u] : 0.f;
          if (g == 0) {
            S.success[prob] = 0;
            for (int r = 0; r < 6; ++r) S.err[prob * 6 + r] = 1e30f;
            if (S.iters) S.iters[prob] = -1;
          }
        } else {
          // set up the problem: oMdes = target * link2joint^-1, q0, mask
          const ModelTable<float> *M = S.tab;
          const int nv = M->nv;
          st.prob = prob;
          st.seg = s;
          st.nc = P.nchain;
          st.it = CONT ? rec->it : 0;
          const LinkRec<float> &Lk = links_of(M)[P.link];
          float Rt[9];
          V3<float> pt;
          pose_to_mat(S.target + tgt * 7, Rt, pt);
#pragma unroll
          for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j)
              st.Rd[3 * i + j] =
                  Rt[3 * i] * Lk.R[3 * j] + Rt[3 * i + 1] * Lk.R[3 * j + 1] + Rt[3 * i + 2] * Lk.R[3 * j + 2];
          st.pd = pt - mv(st.Rd, load3(Lk.p));
          const float *q0 = S.q0 ? S.q0 + prob * nv : nullptr;
#pragma unroll
          for (int c = 0; c < C; ++c) {
            const int j = g * C + c;
            const bool act = j < st.nc;
            const int vu = s_chain[s][act ? j : 0].vuser;
            if constexpr (CONT) {
              st.q[c] = act ? rec->q[j] : 0.f;
              st.qbest[c] = act ? rec->qbest[j] : 0.f;
            } else {
              st.q[c] = act && q0 ? q0[vu] : 0.f;
              st.qbest[c] = st.q[c];
            }
            st.msk[c] = act ? (P.mask ? float(P.mask[prob * P.mask_ld + vu]) : 1.f) : 0.f;
          }
          st.minErr = CONT ? rec->minErr : 1e10f;
          st.midErr = 0.f;
#pragma unroll
          for (int r = 0; r < 6; ++r) st.bestErr[r] = CONT ? rec->bestErr[r] : 0.f;
          state = kRunning;
        }
      }
    }
    if (__ballot(state != kRetired) == 0) break;  // uniform: queue dry and every group done

    // ---- one iteration for every group (only running groups keep the result)
    const ClikSegment &S = s_seg[st.seg];
    const ClikParams &P = S.P;
    const ChainRec *chain = s_chain[st.seg];
    Xf W[C];
    Xf part = identity_xf();
#pragma unroll
    for (int c = 0; c < C; ++c) {
      const int j = g * C + c;
      W[c] = local_xf(chain[j < st.nc ? j : kClikMaxChain], st.q[c]);  // q = 0 past the chain end
      part = c == 0 ? W[0] : compose(part, W[c]);
    }
    Xf incl = part;
    scan_xf<G>(g, incl);
    const Xf excl = sel_xf(g >= 1, below_xf<1>(incl), identity_xf());
    // world transform of each of the lane's joints
#pragma unroll
    for (int c = 0; c < C; ++c) W[c] = compose(c == 0 ? excl : W[c - 1], W[c]);
    const int eel = st.nc > 0 ? (st.nc - 1) / C : 0;
    Xf E = bperm_xf(base4 + 4 * eel, incl);  // the link joint's world pose
    if (st.nc == 0) E = identity_xf();

    // iMd = oMj^-1 oMdes and its log
    float Ri[9];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j)
        Ri[3 * i + j] = E.R[i] * st.Rd[j] + E.R[3 + i] * st.Rd[3 + j] + E.R[6 + i] * st.Rd[6 + j];
    const V3<float> pi = mtv(E.R, st.pd - E.p);
    // (with Jlog6(iMd^-1) for the step from the same rotation log, as its blocks)
    float Ja[9], Jc[9];
    const S6<float> e = log6_jlog6_inv_blocks(Ri, pi, Ja, Jc);
    float ev[6];
    vec_of(e, ev);
    const float en = sqrtf(dot6(e, e));

    // the step, computed by every lane: this lane's columns of
    // Jlog6(iMd^-1) * J_local (masked), J J^T + damp I, its solve in FP64
    // (cond ~1e6+ near singular poses)
    float Jk[C][6];
    double part21[21];
#pragma unroll
    for (int r = 0; r < 21; ++r) part21[r] = 0.0;
#pragma unroll
    for (int c = 0; c < C; ++c) {
      const int j = g * C + c;
      const ChainRec &J = chain[j < st.nc ? j : kClikMaxChain];  // axis 0 past the chain end: column 0
      const V3<float> aw = mv(W[c].R, load3(J.axis));
      const S6<float> cw = J.kind == KIND_PRISMATIC ? S6<float>{aw, {0.f, 0.f, 0.f}} : S6<float>{cross(W[c].p, aw), aw};
      const S6<float> cl = to_local(E.R, E.p, cw);
      const V3<float> u = mv(Ja, cl.a), top = mv(Ja, cl.l) + mv(Jc, u);
      const float jv[6] = {top.x, top.y, top.z, u.x, u.y, u.z};
#pragma unroll
      for (int r = 0; r < 6; ++r) Jk[c][r] = -jv[r] * st.msk[c];
      int t = 0;
#pragma unroll
      for (int r = 0; r < 6; ++r)
#pragma unroll
        for (int m = 0; m <= r; ++m) part21[t++] += double(Jk[c][r]) * double(Jk[c][m]);
    }
    double JJt[36], y[6];
    {
      int t = 0;
#pragma unroll
      for (int r = 0; r < 6; ++r)
#pragma unroll
        for (int m = 0; m <= r; ++m) JJt[r * 6 + m] = group_allsum<G>(part21[t++]);
    }
#pragma unroll
    for (int r = 0; r < 6; ++r) {
      JJt[r * 6 + r] += P.damp;
#pragma unroll
      for (int m = 0; m < r; ++m) JJt[m * 6 + r] = JJt[r * 6 + m];
      y[r] = double(ev[r]);
    }
    ldlt6_solve(JJt, y);

    // first phase: a problem that takes another step and reaches A.cap
    // iterations with it is handed over (slots taken like the refill, in code
    // every lane runs)
    bool cont = false;
    int64_t slot = 0;
    if constexpr (!CONT) {
      const bool fin = en < float(P.eps) || st.it >= P.max_iter;
      cont = state == kRunning && !fin && st.it + 1 == A.cap;
      // A problem on its way converges geometrically (about 0.9 per iteration
      // at dt = 0.1, 0.04 over the second half of the phase); one whose error
      // fell by less than 5x over that half is likely stuck, runs to the
      // iteration limit, and goes to the front.
      const bool far = en > 0.2f * st.midErr;
      const uint64_t wh = __ballot(cont && far && g == 0), we = __ballot(cont && !far && g == 0);
      int base_h = 0, base_e = 0;  // read from lane 0 with every lane active
      if (wh) {                    // uniform
        const int old = atomicAdd(A.cont_count, lane == 0 ? __popcll(wh) : 0);
        base_h = __builtin_amdgcn_readlane(old, 0);
      }
      if (we) {  // uniform
        const int old = atomicAdd(A.cont_easy, lane == 0 ? __popcll(we) : 0);
        base_e = __builtin_amdgcn_readlane(old, 0);
      }
      slot = far ? (int64_t)base_h + __popcll(wh & below) : A.total - 1 - ((int64_t)base_e + __popcll(we & below));
    }

    if (state == kRunning) {  // per group; no cross-lane operations inside
      if (!CONT && st.it == A.cap / 2) st.midErr = en;
      if (en < st.minErr) {
        st.minErr = en;
#pragma unroll
        for (int c = 0; c < C; ++c) st.qbest[c] = st.q[c];
#pragma unroll
        for (int r = 0; r < 6; ++r) st.bestErr[r] = ev[r];
      }
      const bool ok = en < float(P.eps);
      if (ok || st.it >= P.max_iter) {
        // result: chain dofs from the best iterate, the rest from q0; every
        // continuous dof wrapped by atan2 (posP2S over the whole configuration,
        // pinocchio_model.cpp:87,299)
        const int nv = S.tab->nv;
        float *qo = S.q_out + st.prob * nv;
        const float *q0 = S.q0 ? S.q0 + st.prob * nv : nullptr;
        const JointRec<float> *JT = joints_of(S.tab);
        for (int j = g; j < S.tab->nj; j += G) {
          const int u = JT[j].vuser;
          bool chain_dof = false;
          for (int k = 0; k < st.nc; ++k) chain_dof |= chain[k].vuser == u;
          if (!chain_dof) {
            float v = q0 ? q0[u] : 0.f;
            if (JT[j].kind == KIND_CONTINUOUS) {
              float s, co;
              sincosf(v, &s, &co);
              v = atan2f(s, co);
            }
            qo[u] = v;
          }
        }
#pragma unroll
        for (int c = 0; c < C; ++c) {
          const int j = g * C + c;
          if (j < st.nc) {
            float v = st.qbest[c];
            if (chain[j].kind == KIND_CONTINUOUS) {
              float s, co;
              sincosf(v, &s, &co);
              v = atan2f(s, co);
            }
            qo[chain[j].vuser] = v;
          }
        }
        if (g == 0) {
          S.success[st.prob] = ok ? 1 : 0;
#pragma unroll
          for (int r = 0; r < 6; ++r) S.err[st.prob * 6 + r] = st.bestErr[r];
          if (S.iters) S.iters[st.prob] = st.it;
          if (ok && P.solved) {
            const int64_t tgt = P.ntargets ? st.prob % P.ntargets : st.prob;
            const int32_t start = P.ntargets ? (int32_t)(st.prob / P.ntargets) : 0;
            __hip_atomic_fetch_min(P.solved + tgt, start, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          }
        }
        state = kIdle;
      } else {
#pragma unroll
        for (int c = 0; c < C; ++c) {
          float v = 0.f;
#pragma unroll
          for (int r = 0; r < 6; ++r) v -= Jk[c][r] * float(y[r]);
          st.q[c] += v * float(P.dt);
        }
        ++st.it;
        if (cont && slot >= 0 && slot < A.total) {  // always in range: at most A.total hand-overs
          ClikCont &R = A.cont[slot];
#pragma unroll
          for (int c = 0; c < C; ++c) {
            R.q[g * C + c] = st.q[c];
            R.qbest[g * C + c] = st.qbest[c];
          }
          if (g == 0) {
            R.prob = S.first + st.prob;
            R.it = st.it;
            R.minErr = st.minErr;
#pragma unroll
            for (int r = 0; r < 6; ++r) R.bestErr[r] = st.bestErr[r];
          }
          state = kIdle;
        }
      }
    }
  }
}

}  // namespace

int clik_lanes_group(int max_chain) { return max_chain <= kClikMaxChain ? 4 : 0; }

bool clik_schedule_valid(const ClikSchedule &sch) {
  const auto ok = [](int g) { return g == 4 || g == 8 || g == 16; };
  return sch.cap >= 0 && ok(sch.g1) && (sch.cap == 0 || ok(sch.g2));
}

hipError_t launch_clik_lanes(const HeteroClikArgs &A, int max_chain, int cus, const ClikSchedule &sch,
                             hipStream_t s) {
  if (A.total <= 0) return hipSuccess;
  if (!clik_lanes_group(max_chain) || A.total > INT32_MAX || !A.cont || !A.cont_count || !A.cont_queue ||
      !A.cont_easy || !clik_schedule_valid(sch))
    return hipErrorNotSupported;
  // The same schedule for every chain length, so a problem gives the same bits
  // in a per-model launch and in a mixed-tree launch. By default the first
  // phase packs 16 problems per wavefront (throughput); the problems left
  // after its cap of iterations, the slow and the unreachable ones, run 8 per
  // wavefront with half the joints per lane (a shorter iteration).
  // Persistent grids, fewer blocks for small batches: the first phase two
  // blocks per CU (two wavefronts per SIMD), the second one block per CU, so
  // the long problems it starts first keep a SIMD each.
  HeteroClikArgs A1 = A;
  A1.cap = sch.cap;
  const auto grid = [&](int per_block, int per_cu) {
    return (unsigned)std::max<int64_t>(1, std::min<int64_t>(per_cu * cus, (A.total + per_block - 1) / per_block));
  };
  switch (sch.g1) {
    case 4: hipLaunchKernelGGL((k_clik_lanes<4, 4, false>), dim3(grid(256 / 4, 2)), dim3(256), 0, s, A1); break;
    case 8: hipLaunchKernelGGL((k_clik_lanes<8, 2, false>), dim3(grid(256 / 8, 2)), dim3(256), 0, s, A1); break;
    default: hipLaunchKernelGGL((k_clik_lanes<16, 1, false>), dim3(grid(256 / 16, 2)), dim3(256), 0, s, A1); break;
  }
  if (sch.cap > 0) {
    switch (sch.g2) {
      case 4: hipLaunchKernelGGL((k_clik_lanes<4, 4, true>), dim3(grid(256 / 4, 1)), dim3(256), 0, s, A1); break;
      case 8: hipLaunchKernelGGL((k_clik_lanes<8, 2, true>), dim3(grid(256 / 8, 1)), dim3(256), 0, s, A1); break;
      default: hipLaunchKernelGGL((k_clik_lanes<16, 1, true>), dim3(grid(256 / 16, 1)), dim3(256), 0, s, A1); break;
    }
  }
  return hipGetLastError();
}

}  // namespace sapamd
