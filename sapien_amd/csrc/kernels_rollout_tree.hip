// Lanes-per-joint forward dynamics and rollout steps for trees of <= 32
// joints: one environment per group of G lanes, lane j owning joint j,
// everything in registers and LDS (no private segment). The arithmetic type
// Real is float, or double for ill-conditioned mass matrices; inputs, outputs
// and the rollout state stay FP32 either way.
//
// Forward dynamics by factorization instead of ABA's three sweeps:
//   b   = RNEA(q, qd, 0)              tree prefixes (v, a) and subtree sums (f)
//   M   = CRBA(q)                     subtree sums of the inertias; lane j
//                                     writes row j of M into the group's LDS
//   qdd = M^-1 (tau - b)              LDL^T of M in LDS, lane i updating row i
// A rollout step is qd' = qd + dt qdd, q' = q + dt qd'. The step adjoint reuses
// the factorization for lambda = M^-1 (dt c) and the lanes-per-joint RNEA
// vector-Jacobian product (the formulas of kernels_adjoint.hip):
//   c = abar_qd' + dt abar_q',  dL/dtau = lambda,
//   dL/dq = abar_q' - (dRNEA/dq)^T lambda,  dL/dqd = c - (dRNEA/dqd)^T lambda.
// All lanes of a group are in one wavefront, so the LDS factorization needs
// only wavefront-level ordering (no block barrier).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <type_traits>

#include "lanes.h"
#include "launch.h"
#include "spatial.h"
#include "tree.h"

namespace sapamd {

namespace {

using namespace lanes;

constexpr int kTreeDynJoints = 32;
template <class Real>
constexpr int kDynTableWords4 = (int)((kHeaderBytes + sizeof(JointRec<Real>) * kTreeDynJoints) / 16);


// Compile-time loop: f(std::integral_constant<int, K>) for K in [B, E), so
// register arrays indexed by K never fall back to the private segment (the
// unroller's size limit does not apply).
// The lane's index within its group, opaque to the optimizer: the
// factorization compares it with every column index, and inside a time-step
// loop those loop-invariant compares would otherwise be hoisted out of the
// loop and spilled (a compare is one instruction, a spill round trip two).
__device__ __forceinline__ int opaque_lane(int g) {
  asm volatile("" : "+v"(g));
  return g;
}

// 1 / x by the hardware reciprocal refined with Newton-Raphson steps
// r' = r + r (1 - x r) (one for FP32, two for FP64): within an ulp or two of
// the quotient, without the full IEEE division sequence on the critical path of
// every elimination step.
__device__ __forceinline__ float recip(float x) {
  const float r = __builtin_amdgcn_rcpf(x);
  return __builtin_fmaf(r, __builtin_fmaf(-x, r, 1.0f), r);
}
__device__ __forceinline__ double recip(double x) {
  double r = __builtin_amdgcn_rcp(x);
  r = __builtin_fma(r, __builtin_fma(-x, r, 1.0), r);
  return __builtin_fma(r, __builtin_fma(-x, r, 1.0), r);
}

template <int B, int E, class F> __device__ __forceinline__ void static_for(F &&f) {
  if constexpr (B < E) {
    f(std::integral_constant<int, B>{});
    static_for<B + 1, E>(f);
  }
}

// LDL^T of one diagonal block of the symmetric positive definite M (M is block
// diagonal over the root subtrees of the tree; with one block, M itself). The
// block's n rows sit on lanes o .. o + n - 1 of the group; lane o + i holds
// block row i (lower triangle, zeros above the diagonal, block column j at
// row[j]). Right-looking elimination with the rows in registers, step k
// broadcasting block column k through the group's G-entry LDS buffer (indexed
// by lane; lane is this lane's index in the group). After the call
// row[j < i] = L_ij (entries above the diagonal hold junk: the update runs
// over the whole row to need no per-entry select, and nothing reads them);
// returns D_i. Lanes with i >= n hold zero rows. The steps run over the first
// NMAX >= n columns, a compile-time bound (steps past n change nothing), so no
// step depends on a runtime value; o + NMAX <= G keeps every read in the group.
template <int G, int NMAX, class Real, bool BLOCKS>
__device__ __forceinline__ Real ldlt_factor(Real (&row)[G], Real *buf, int lane, int o, int i, int n) {
  Real d = Real(1);
  const Real *blk = BLOCKS ? buf + o : buf;
  static_for<0, NMAX>([&](auto kc) {
    constexpr int k = decltype(kc)::value;
    buf[lane] = row[k];
    wave_sync();
    Real c[G];
    static_for<k, NMAX>([&](auto jc) { c[decltype(jc)::value] = blk[decltype(jc)::value]; });
    wave_sync();
    const bool act = i > k && i < n;
    const Real l = act ? row[k] * recip(c[k]) : Real(0);
    static_for<k + 1, NMAX>([&](auto jc) {
      constexpr int j = decltype(jc)::value;
      row[j] -= l * c[j];
    });
    row[k] = act ? l : row[k];
    d = i == k ? row[k] : d;
  });
  return d;
}

// R right-hand sides r[m] (block row i's entry on lane o + i, zero for lanes
// with i >= n) -> (L D L^T)^-1 r over the block. Forward substitution reads L
// from the lane's row; the rows are then transposed through the group's LDS
// tile so backward substitution reads lane i's block column. Each step
// broadcasts the finished entry of block row k (lane o + k) with ds_bpermute;
// the factor entries are masked once (strictly lower row, strictly
// below-diagonal column within the block), so the steps need no selects.
template <int G, int NMAX, class Real, int R, bool BLOCKS>
__device__ __forceinline__ void ldlt_solve(const Real (&row)[G], Real d, Real (*A)[G + 1], int base4, int lane, int o,
                                           int i, int n, Real (&r)[R]) {
  const int ob4 = BLOCKS ? base4 + 4 * o : base4;
  Real lo[G];
  static_for<0, NMAX>([&](auto kc) {
    constexpr int k = decltype(kc)::value;
    lo[k] = k < i ? row[k] : Real(0);
  });
  static_for<0, NMAX>([&](auto kc) {
    constexpr int k = decltype(kc)::value;
#pragma unroll
    for (int m = 0; m < R; ++m) r[m] -= lo[k] * bperm(ob4 + 4 * k, r[m]);
  });
  const Real dinv = recip(d);
#pragma unroll
  for (int m = 0; m < R; ++m) r[m] = i < n ? r[m] * dinv : Real(0);
  static_for<0, NMAX>([&](auto jc) { A[lane][decltype(jc)::value] = row[decltype(jc)::value]; });
  wave_sync();
  Real col[G];
  const Real(*Ab)[G + 1] = BLOCKS ? A + o : A;
  static_for<0, NMAX>([&](auto kc) {
    constexpr int k = decltype(kc)::value;
    // rows past the block belong to the next blocks; a single block is followed by zero rows
    col[k] = (BLOCKS ? k > i && k < n : k > i) ? Ab[k][i] : Real(0);
  });
  wave_sync();
  static_for<0, NMAX>([&](auto kc) {
    constexpr int k = NMAX - 1 - decltype(kc)::value;
#pragma unroll
    for (int m = 0; m < R; ++m) r[m] -= col[k] * bperm(ob4 + 4 * k, r[m]);
  });
}

// Per group: the transposition tile (rows padded by one word, so lanes
// walking a row or a column hit different banks) and the column buffer.
template <int G, class Real> struct DynLds {
  Real A[G][G + 1];
  Real y[G];
  Real S[G][6];  // each joint's world column, for the rows of M
};

// One environment's kinematics and M / bias; the factorization is left in L.
template <int G, class Real> struct TreeFd {
  Tree<G, Real> T;
  int g, nj, depth;
  int o, nb;  // first lane and size of this joint's block of M (its root subtree, or the whole tree)
  bool valid;
  const JointRec<Real> *J;
  XfT<Real> X;
  S6<Real> col, v;

  // q, qd in joint order (this lane's joint); loads row with the lower
  // triangle of row g of M (zeros above the diagonal), assembled in the
  // group's LDS tile (block columns: column i at i - o), and returns the bias
  // b_g = RNEA(q, qd, 0)_g.
  __device__ __forceinline__ Real setup(const ModelTable<Real> *M, Real qj, Real qdj, DynLds<G, Real> &L,
                                        Real (&row)[G]) {
    Real s, c;
    joint_sincos(qj, &s, &c);
    joint_world(*J, qj, s, c, static_cast<const Real *>(nullptr), V3<Real>{0, 0, 0}, true, X.R, X.p);
    if (!valid) X = identity_xf<Real>();
    T.prefix(X);
    col = valid ? jcol(*J, X.R, X.p) : zero6<Real>();
    v = qdj * col;
    T.prefix(v);
    S6<Real> a = qdj * crossm(v, col);
    T.prefix(a);
    a.l = a.l - V3<Real>{M->gravity[0], M->gravity[1], M->gravity[2]};
    const WInertia<Real> Y = world_inertia_fast(*J, X.R, X.p);
    S6<Real> f = valid ? inertia_apply(Y, a) + crossf(v, inertia_apply(Y, v)) : zero6<Real>();
    T.subtree(f);
    const Real bias = dot6(col, f);
    // composite inertia of the subtree and the lower triangle of row g of M
    AInertia<Real> Ya = additive(Y);
    Real y[10] = {Ya.m, Ya.h.x, Ya.h.y, Ya.h.z, Ya.I[0], Ya.I[1], Ya.I[2], Ya.I[3], Ya.I[4], Ya.I[5]};
    if (!valid)
#pragma unroll
      for (int k = 0; k < 10; ++k) y[k] = Real(0);
    T.template subtree<10>(y);
    AInertia<Real> Yc;
    Yc.m = y[0];
    Yc.h = {y[1], y[2], y[3]};
#pragma unroll
    for (int k = 0; k < 6; ++k) Yc.I[k] = y[4 + k];
    const S6<Real> Fg = apply(Yc, col);
    // every lane publishes its column; lane g reads those of its ancestors
    // (one address across the group at a time: broadcast reads, no chain of
    // permutes)
    L.S[g][0] = col.l.x;
    L.S[g][1] = col.l.y;
    L.S[g][2] = col.l.z;
    L.S[g][3] = col.a.x;
    L.S[g][4] = col.a.y;
    L.S[g][5] = col.a.z;
#pragma unroll
    for (int i = 0; i < G; ++i) L.A[g][i] = Real(0);
    wave_sync();
    // M_gi is nonzero only for i an ancestor of g (or g): walk the set bits of
    // the ancestor mask, at most the tree depth of them, instead of every joint
    uint32_t am = valid ? J->ancmask : 0u;
    for (int k = 0; k < depth; ++k) {
      const int i = am ? __builtin_ctz(am) : 0;
      const S6<Real> ci = {{L.S[i][0], L.S[i][1], L.S[i][2]}, {L.S[i][3], L.S[i][4], L.S[i][5]}};
      const Real dij = dot6(ci, Fg);
      if (am) L.A[g][i - o] = dij;
      am &= am - 1u;
    }
    wave_sync();
    static_for<0, G>([&](auto ic) { row[decltype(ic)::value] = L.A[g][decltype(ic)::value]; });
    wave_sync();
    return bias;
  }
};

template <int G, class Real, bool BLOCKS>
__device__ __forceinline__ void init_fd(TreeFd<G, Real> &F, const ModelTable<Real> *M, int g) {
  F.g = g;
  F.nj = __builtin_amdgcn_readfirstlane(M->nj);  // the same in every lane: a scalar
  F.depth = __builtin_amdgcn_readfirstlane(M->max_depth);
  F.valid = g < F.nj;
  F.J = joints_of(M) + (F.valid ? g : F.nj - 1);
  F.T.g = g;
  F.T.base4 = ((threadIdx.x & 63) & ~(G - 1)) * 4;
#pragma unroll
  for (int k = 0; k < Tree<G, Real>::K; ++k) F.T.anc[k] = F.valid ? F.J->anc[k] : -1;
  F.T.size = F.valid ? F.J->subtree : 0;
  if constexpr (BLOCKS) {  // per root subtree: the lowest ancestor bit is its root
    F.o = F.valid ? (int)__builtin_ctz(F.J->ancmask) : 0;
    F.nb = F.valid ? joints_of(M)[F.o].subtree : 0;
  } else {
    F.o = 0;
    F.nb = F.nj;
  }
}

// Stage the model table in LDS (block-wide) and return the LDS copy.
template <class Real>
__device__ __forceinline__ const ModelTable<Real> *stage_table(const ModelTable<Real> *Mg, float4 *s_tab) {
  const int nj = Mg->nj;
  const int n4 = (int)((kHeaderBytes + sizeof(JointRec<Real>) * nj) / 16);
  const float4 *src = reinterpret_cast<const float4 *>(Mg);
  for (int i = threadIdx.x; i < n4; i += blockDim.x) s_tab[i] = src[i];
  __syncthreads();
  return reinterpret_cast<const ModelTable<Real> *>(s_tab);
}

// Forward dynamics of this lane's joint at (q, qd, tau) (joint order).
template <int G, int NMAX, class Real, bool BLOCKS>
__device__ __forceinline__ Real lane_fd(TreeFd<G, Real> &F, const ModelTable<Real> *M, DynLds<G, Real> &L, Real qj,
                                        Real qdj, Real tj) {
  Real row[G];
  const Real bias = F.setup(M, qj, qdj, L, row);
  const int i = opaque_lane(F.g - F.o);
  const Real d = ldlt_factor<G, NMAX, Real, BLOCKS>(row, L.y, F.g, F.o, i, F.nb);
  Real r[1] = {tj - bias};
  ldlt_solve<G, NMAX, Real, 1, BLOCKS>(row, d, L.A, F.T.base4, F.g, F.o, i, F.nb, r);
  return r[0];
}

// Adjoint of one semi-implicit Euler step at the pre-step state (q, qd) for
// this lane's joint: (aq, aqd) post-step adjoint in, pre-step adjoint out;
// returns dL/dtau_j = lambda_j.
template <int G, int NMAX, class Real, bool BLOCKS>
__device__ __forceinline__ Real lane_step_vjp(TreeFd<G, Real> &F, const ModelTable<Real> *M, DynLds<G, Real> &L,
                                              float dt, Real qj, Real qdj, Real tj, Real &aqj, Real &aqdj) {
  Real row[G];
  const Real bias = F.setup(M, qj, qdj, L, row);
  const int i = opaque_lane(F.g - F.o);
  const Real d = ldlt_factor<G, NMAX, Real, BLOCKS>(row, L.y, F.g, F.o, i, F.nb);
  const Real dtr = dt;
  const Real cj = aqdj + dtr * aqj;
  Real r[2] = {tj - bias, dtr * cj};  // qdd and lambda = M^-1 (dt c) in one pass
  ldlt_solve<G, NMAX, Real, 2, BLOCKS>(row, d, L.A, F.T.base4, F.g, F.o, i, F.nb, r);
  const Real qddj = r[0], lam = r[1];

  // RNEA vector-Jacobian product at (q, qd, qdd) for the upstream lambda
  const Tree<G, Real> &T = F.T;
  const S6<Real> col = F.col, v = F.v;
  S6<Real> a = qddj * col + qdj * crossm(v, col);
  T.prefix(a);
  a.l = a.l - V3<Real>{M->gravity[0], M->gravity[1], M->gravity[2]};
  const WInertia<Real> Y = world_inertia_fast(*F.J, F.X.R, F.X.p);
  S6<Real> Ff = F.valid ? inertia_apply(Y, a) + crossf(v, inertia_apply(Y, v)) : zero6<Real>();
  T.subtree(Ff);
  const Real gj = F.valid ? lam : Real(0);
  S6<Real> w = gj * col;
  T.prefix(w);
  const S6<Real> Iw = inertia_apply(Y, w), Iv = inertia_apply(Y, v);
  S6<Real> Hw = zero6<Real>(), QS = zero6<Real>();
  if (F.valid) {
    Hw = Iw;
    QS = crossf(w, Iv) - crossf(v, Iw) + inertia_apply(Y, crossm(v, w));
  }
  Real acc[12] = {Hw.l.x, Hw.l.y, Hw.l.z, Hw.a.x, Hw.a.y, Hw.a.z, QS.l.x, QS.l.y, QS.l.z, QS.a.x, QS.a.y, QS.a.z};
  T.template subtree<12>(acc);
  Hw = {{acc[0], acc[1], acc[2]}, {acc[3], acc[4], acc[5]}};
  QS = {{acc[6], acc[7], acc[8]}, {acc[9], acc[10], acc[11]}};
  const S6<Real> wl = w - gj * col, vl = v - qdj * col;
  const S6<Real> al = a - qddj * col - qdj * crossm(v, col);
  const Real dq = dot6(col, crossf(vl, QS + crossf(vl, Hw)) - crossf(wl, Ff) - crossf(al, Hw));
  const Real dqd = -dot6(col, QS + Real(2) * crossf(vl, Hw));
  aqj = aqj - dq;
  aqdj = cj - dqd;
  return lam;
}

// Shared prologue of the kernels: group lane, environment (clamped for the
// tail groups), staged table and this lane's joint.
#define SAPAMD_TREE_DYN_PROLOGUE                                           \
  __shared__ float4 s_tab[kDynTableWords4<Real>];                          \
  __shared__ DynLds<G, Real> s_dyn[BLOCK / G];                             \
  const int g = threadIdx.x & (G - 1);                                     \
  const int64_t env = (blockIdx.x * (int64_t)BLOCK + threadIdx.x) / G;     \
  const bool live = env < B;                                               \
  const int64_t e = live ? env : B - 1;                                    \
  const int nv = Mg->nv;

// mode 0: forward dynamics qdd = FD(q, qd, tau) into out (B, nv)
// mode 1: rollout step in place on q, qd; save (pre-step q | qd) when given
template <int G, class Real, int BLOCK, int NMAX, bool BLOCKS>
__global__ __launch_bounds__(BLOCK) void k_tree_step(const ModelTable<Real> *__restrict__ Mg, int64_t B, int mode,
                                                     float dt, float *__restrict__ q, float *__restrict__ qd,
                                                     const float *__restrict__ tau, int64_t tau_stride,
                                                     float *__restrict__ out, float *__restrict__ save,
                                                     int64_t save_stride) {
  SAPAMD_TREE_DYN_PROLOGUE
  // loads first (one round trip): lane g reads user dof g
  const int gc = g < nv ? g : nv - 1;
  const float qu = q[e * nv + gc], qdu = qd[e * nv + gc], tu = tau[e * tau_stride + gc];
  const ModelTable<Real> *M = stage_table(Mg, s_tab);
  DynLds<G, Real> &L = s_dyn[threadIdx.x / G];
  TreeFd<G, Real> F;
  init_fd<G, Real, BLOCKS>(F, M, g);
  const int src = F.T.base4 + 4 * F.J->vuser;
  Real qj = bperm(src, qu), qdj = bperm(src, qdu), tj = bperm(src, tu);
  if (!F.valid) qj = qdj = tj = Real(0);
  const Real qddj = lane_fd<G, NMAX, Real, BLOCKS>(F, M, L, qj, qdj, tj);
  if (!live) return;
  if (mode == 0) {
    if (F.valid) out[e * nv + F.J->vuser] = (float)qddj;
    return;
  }
  if (save && g < nv) {
    save[e * save_stride + g] = qu;
    save[e * save_stride + nv + g] = qdu;
  }
  if (F.valid) {
    // semi-implicit Euler on the FP32 state
    const float v1 = (float)qdj + dt * (float)qddj;
    qd[e * nv + F.J->vuser] = v1;
    q[e * nv + F.J->vuser] = (float)qj + dt * v1;
  }
}

// Adjoint of one step; state = pre-step (q | qd) rows, aq / aqd hold the
// post-step adjoint on entry and the pre-step adjoint on exit.
template <int G, class Real, int BLOCK, int NMAX, bool BLOCKS>
__global__ __launch_bounds__(BLOCK) void k_tree_step_vjp(const ModelTable<Real> *__restrict__ Mg, int64_t B, float dt,
                                                         const float *__restrict__ state, int64_t state_stride,
                                                         const float *__restrict__ tau, int64_t tau_stride,
                                                         float *__restrict__ aq, float *__restrict__ aqd,
                                                         float *__restrict__ gtau) {
  SAPAMD_TREE_DYN_PROLOGUE
  const int gc = g < nv ? g : nv - 1;
  const float qu = state[e * state_stride + gc], qdu = state[e * state_stride + nv + gc];
  const float tu = tau[e * tau_stride + gc], aqu = aq[e * nv + gc], aqdu = aqd[e * nv + gc];
  const ModelTable<Real> *M = stage_table(Mg, s_tab);
  DynLds<G, Real> &L = s_dyn[threadIdx.x / G];
  TreeFd<G, Real> F;
  init_fd<G, Real, BLOCKS>(F, M, g);
  const int src = F.T.base4 + 4 * F.J->vuser;
  Real qj = bperm(src, qu), qdj = bperm(src, qdu), tj = bperm(src, tu);
  Real aqj = bperm(src, aqu), aqdj = bperm(src, aqdu);
  if (!F.valid) qj = qdj = tj = aqj = aqdj = Real(0);
  const Real lam = lane_step_vjp<G, NMAX, Real, BLOCKS>(F, M, L, dt, qj, qdj, tj, aqj, aqdj);
  if (!live || !F.valid) return;
  const int u = F.J->vuser;
  aq[e * nv + u] = (float)aqj;
  aqd[e * nv + u] = (float)aqdj;
  gtau[e * tau_stride + u] = (float)lam;
}

// T steps of semi-implicit Euler for environment env of a batch of B, the
// state held in registers (lane j: q_j, qd_j in FP32, exactly as the one-step
// kernel rounds it). M is the model table staged in LDS, L the group's tile.
// tau (B, T, nv); ckpt (B, ceil(T / K), 2 nv) receives the pre-step state of
// steps 0, K, 2K, ... when given; q, qd (B, nv) hold the initial state on entry
// and the final state on exit.
template <int G, class Real, int NMAX, bool BLOCKS>
__device__ __forceinline__ void rollout_env(const ModelTable<Real> *M, DynLds<G, Real> &L, int64_t env, int64_t B,
                                            int T, int K, float dt, float *__restrict__ q, float *__restrict__ qd,
                                            const float *__restrict__ tau, float *__restrict__ ckpt) {
  const int g = threadIdx.x & (G - 1);
  const bool live = env < B;
  const int64_t e = live ? env : B - 1;  // tail groups recompute the last environment and store nothing
  const int nv = M->nv;
  TreeFd<G, Real> F;
  init_fd<G, Real, BLOCKS>(F, M, g);
  const int u = F.J->vuser;  // lanes past the last joint repeat its dof and store nothing
  const bool owner = live && F.valid;
  const int64_t C = (T + K - 1) / K;
  const float *tr = tau + e * (int64_t)T * nv + u;
  float qs = q[e * nv + u], qds = qd[e * nv + u];
  float tnext = tr[0];
  for (int t = 0; t < T; ++t) {
    const float tcur = tnext;
    if (t + 1 < T) tnext = tr[(int64_t)(t + 1) * nv];
    if (ckpt && owner && t % K == 0) {
      float *c = ckpt + (e * C + t / K) * 2 * nv;
      c[u] = qs;
      c[nv + u] = qds;
    }
    const Real qddj = lane_fd<G, NMAX, Real, BLOCKS>(F, M, L, F.valid ? Real(qs) : Real(0), F.valid ? Real(qds) : Real(0),
                                       F.valid ? Real(tcur) : Real(0));
    const float v1 = qds + dt * (float)qddj;
    qs = qs + dt * v1;
    qds = v1;
  }
  if (owner) {
    q[e * nv + u] = qs;
    qd[e * nv + u] = qds;
  }
}

// Adjoint of rollout_env for environment env: segments of K steps from the
// last, each recomputed from its checkpoint with the pre-step states kept in
// seg (B, K, 2 nv; each lane reads back only what it wrote), then the step
// adjoints in reverse. aq, aqd: dL/d(final state) in, dL/d(initial state)
// out; gtau (B, T, nv) = dL/dtau.
template <int G, class Real, int NMAX, bool BLOCKS>
__device__ __forceinline__ void rollout_vjp_env(const ModelTable<Real> *M, DynLds<G, Real> &L, int64_t env,
                                                int64_t B, int T, int K, float dt, const float *__restrict__ tau,
                                                const float *__restrict__ ckpt, float *__restrict__ seg,
                                                float *__restrict__ aq, float *__restrict__ aqd,
                                                float *__restrict__ gtau) {
  const int g = threadIdx.x & (G - 1);
  const bool live = env < B;
  const int64_t e = live ? env : B - 1;
  const int nv = M->nv;
  TreeFd<G, Real> F;
  init_fd<G, Real, BLOCKS>(F, M, g);
  const int u = F.J->vuser;
  const bool owner = live && F.valid;
  const int C = (T + K - 1) / K;
  const float *tr = tau + e * (int64_t)T * nv + u;
  float *gr = gtau + e * (int64_t)T * nv + u;
  float *sg = seg + e * (int64_t)K * 2 * nv + u;
  Real aqj = F.valid ? Real(aq[e * nv + u]) : Real(0), aqdj = F.valid ? Real(aqd[e * nv + u]) : Real(0);
  for (int c = C - 1; c >= 0; --c) {
    const int t0 = c * K, t1 = min(T, t0 + K);
    const float *ck = ckpt + ((int64_t)e * C + c) * 2 * nv + u;
    float qs = ck[0], qds = ck[nv];
    for (int t = t0; t < t1; ++t) {
      if (owner) {
        sg[(int64_t)(t - t0) * 2 * nv] = qs;
        sg[(int64_t)(t - t0) * 2 * nv + nv] = qds;
      }
      if (t + 1 == t1) break;  // the last pre-step state needs no step after it
      const Real qddj = lane_fd<G, NMAX, Real, BLOCKS>(F, M, L, F.valid ? Real(qs) : Real(0), F.valid ? Real(qds) : Real(0),
                                         F.valid ? Real(tr[(int64_t)t * nv]) : Real(0));
      const float v1 = qds + dt * (float)qddj;
      qs = qs + dt * v1;
      qds = v1;
    }
    for (int t = t1 - 1; t >= t0; --t) {
      const Real qj = F.valid ? Real(sg[(int64_t)(t - t0) * 2 * nv]) : Real(0);
      const Real qdj = F.valid ? Real(sg[(int64_t)(t - t0) * 2 * nv + nv]) : Real(0);
      const Real tj = F.valid ? Real(tr[(int64_t)t * nv]) : Real(0);
      const Real lam = lane_step_vjp<G, NMAX, Real, BLOCKS>(F, M, L, dt, qj, qdj, tj, aqj, aqdj);
      // the one-step kernel hands the adjoint on in FP32: round the same way
      aqj = F.valid ? Real((float)aqj) : Real(0);
      aqdj = F.valid ? Real((float)aqdj) : Real(0);
      if (owner) gr[(int64_t)t * nv] = (float)lam;
    }
  }
  if (owner) {
    aq[e * nv + u] = (float)aqj;
    aqd[e * nv + u] = (float)aqdj;
  }
}

// T steps in one launch, one model: blocks of BLOCK threads, G per environment.
template <int G, class Real, int BLOCK, int NMAX, bool BLOCKS>
__global__ __launch_bounds__(BLOCK) void k_tree_rollout(const ModelTable<Real> *__restrict__ Mg, int64_t B, int T,
                                                        int K, float dt, float *__restrict__ q,
                                                        float *__restrict__ qd, const float *__restrict__ tau,
                                                        float *__restrict__ ckpt) {
  __shared__ float4 s_tab[kDynTableWords4<Real>];
  __shared__ DynLds<G, Real> s_dyn[BLOCK / G];
  const ModelTable<Real> *M = stage_table(Mg, s_tab);
  rollout_env<G, Real, NMAX, BLOCKS>(M, s_dyn[threadIdx.x / G], (blockIdx.x * (int64_t)BLOCK + threadIdx.x) / G, B, T, K,
                             dt, q, qd, tau, ckpt);
}

// Adjoint of k_tree_rollout in one launch.
template <int G, class Real, int BLOCK, int NMAX, bool BLOCKS>
__global__ __launch_bounds__(BLOCK) void k_tree_rollout_vjp(const ModelTable<Real> *__restrict__ Mg, int64_t B, int T,
                                                            int K, float dt, const float *__restrict__ tau,
                                                            const float *__restrict__ ckpt, float *__restrict__ seg,
                                                            float *__restrict__ aq, float *__restrict__ aqd,
                                                            float *__restrict__ gtau) {
  __shared__ float4 s_tab[kDynTableWords4<Real>];
  __shared__ DynLds<G, Real> s_dyn[BLOCK / G];
  const ModelTable<Real> *M = stage_table(Mg, s_tab);
  rollout_vjp_env<G, Real, NMAX, BLOCKS>(M, s_dyn[threadIdx.x / G], (blockIdx.x * (int64_t)BLOCK + threadIdx.x) / G, B, T,
                                 K, dt, tau, ckpt, seg, aq, aqd, gtau);
}

// Rollouts of several models in one launch: block ranges map to segments (one
// model each, ordered by the host); the table and the groups' tiles are carved
// from one dynamic LDS allocation sized for the largest segment. Every case
// of the switch is the per-model kernels' code, so each environment gets the
// same bits as in its own model's launch.
constexpr int kMixBlock = 128;

template <int G, class Real, int NMAX, bool BLOCKS, bool VJP>
__device__ __forceinline__ void mixed_segment(const RolloutSegment &S, const RolloutMixArgs &A, float4 *lds) {
  const ModelTable<Real> *M = stage_table(static_cast<const ModelTable<Real> *>(S.tab), lds);
  DynLds<G, Real> *dyn = reinterpret_cast<DynLds<G, Real> *>(lds + kDynTableWords4<Real>);
  const int64_t env = ((int64_t)(blockIdx.x - S.block0) * kMixBlock + threadIdx.x) / G;
  if constexpr (VJP)
    rollout_vjp_env<G, Real, NMAX, BLOCKS>(M, dyn[threadIdx.x / G], env, S.B, A.T, A.K, A.dt, S.tau, S.ckpt, S.seg, S.aq,
                                   S.aqd, S.gtau);
  else
    rollout_env<G, Real, NMAX, BLOCKS>(M, dyn[threadIdx.x / G], env, S.B, A.T, A.K, A.dt, S.q, S.qd, S.tau, S.ckpt_out);
}

template <bool VJP> __global__ __launch_bounds__(kMixBlock) void k_tree_rollout_mixed(RolloutMixArgs A) {
  extern __shared__ float4 s_mix[];
  int s = 0;
  while (s + 1 < A.nseg && (int64_t)blockIdx.x >= A.seg[s + 1].block0) ++s;
  const RolloutSegment &S = A.seg[s];
  switch (S.variant) {  // uniform per block
    case 0: mixed_segment<8, float, 8, false, VJP>(S, A, s_mix); break;
    case 1: mixed_segment<16, float, 12, false, VJP>(S, A, s_mix); break;
    case 2: mixed_segment<16, float, 16, false, VJP>(S, A, s_mix); break;
    case 3: mixed_segment<32, float, 20, false, VJP>(S, A, s_mix); break;
    case 4: mixed_segment<32, float, 24, false, VJP>(S, A, s_mix); break;
    case 5: mixed_segment<32, float, 32, false, VJP>(S, A, s_mix); break;
    case 6: mixed_segment<32, float, 20, true, VJP>(S, A, s_mix); break;
    case 7: mixed_segment<32, float, 24, true, VJP>(S, A, s_mix); break;
    case 8: mixed_segment<8, double, 8, false, VJP>(S, A, s_mix); break;
    case 9: mixed_segment<16, double, 12, false, VJP>(S, A, s_mix); break;
    case 10: mixed_segment<16, double, 16, false, VJP>(S, A, s_mix); break;
    case 11: mixed_segment<32, double, 20, false, VJP>(S, A, s_mix); break;
    case 12: mixed_segment<32, double, 24, false, VJP>(S, A, s_mix); break;
    case 13: mixed_segment<32, double, 32, false, VJP>(S, A, s_mix); break;
    case 14: mixed_segment<32, double, 20, true, VJP>(S, A, s_mix); break;
    default: mixed_segment<32, double, 24, true, VJP>(S, A, s_mix); break;
  }
}

// The switch case for nj joints and a factorization of `rows` rows: (G, NMAX)
// as SAPAMD_TREE_DYN_LAUNCH picks them, per-block factorizations (rows < nj)
// after the single-block ones, FP64 cases after the FP32 ones.
constexpr int kMixVariants = 16;
int mixed_variant(int nj, int rows, bool fp64) {
  const int G = tree_dyn_group(nj), ns = tree_dyn_steps(rows, G);
  const int v = G == 8    ? 0
                : G == 16 ? (ns == 12 ? 1 : 2)
                : rows < nj ? (ns == 20 ? 6 : 7)
                            : (ns == 20 ? 3 : ns == 24 ? 4 : 5);
  return v + (fp64 ? kMixVariants / 2 : 0);
}

template <class Real, int G> constexpr size_t mixed_lds() {
  return sizeof(float4) * kDynTableWords4<Real> + sizeof(DynLds<G, Real>) * (kMixBlock / G);
}
static_assert(mixed_lds<double, 32>() <= 65536, "the mixed launch's LDS must fit the default dynamic LDS limit");
size_t mixed_lds_bytes(int variant) {
  constexpr size_t t[kMixVariants] = {
      mixed_lds<float, 8>(),   mixed_lds<float, 16>(),  mixed_lds<float, 16>(),  mixed_lds<float, 32>(),
      mixed_lds<float, 32>(),  mixed_lds<float, 32>(),  mixed_lds<float, 32>(),  mixed_lds<float, 32>(),
      mixed_lds<double, 8>(),  mixed_lds<double, 16>(), mixed_lds<double, 16>(), mixed_lds<double, 32>(),
      mixed_lds<double, 32>(), mixed_lds<double, 32>(), mixed_lds<double, 32>(), mixed_lds<double, 32>()};
  return t[variant];
}

// FP32: 256-thread blocks; FP64: 128-thread blocks (half the LDS per block).
template <class Real> constexpr int kStepBlock = sizeof(Real) == 4 ? 256 : 128;

// Launches KERNEL<G, Real, BL, NMAX, BLOCKS> for nj joints and a
// factorization of nr rows: G lanes per environment, the elimination bounded
// by NMAX = tree_dyn_steps(nr, G), per root subtree when nr < nj (planned by
// dyn_factor_plan only for 32-lane groups and 20 or 24 steps).
#define SAPAMD_TREE_DYN_LAUNCH(KERNEL, ...)                                                          \
  do {                                                                                               \
    constexpr int BL = kStepBlock<Real>;                                                             \
    const int G = tree_dyn_group(nj);                                                                \
    const int NS = tree_dyn_steps(nr, G);                                                            \
    const bool blocks = nr < nj;                                                                     \
    const dim3 grid((unsigned)((B * G + BL - 1) / BL));                                              \
    if (blocks && (G != 32 || (NS != 20 && NS != 24))) return hipErrorNotSupported;                  \
    if (G == 8)                                                                                      \
      hipLaunchKernelGGL((KERNEL<8, Real, BL, 8, false>), grid, dim3(BL), 0, s, __VA_ARGS__);        \
    else if (G == 16 && NS == 12)                                                                    \
      hipLaunchKernelGGL((KERNEL<16, Real, BL, 12, false>), grid, dim3(BL), 0, s, __VA_ARGS__);      \
    else if (G == 16)                                                                                \
      hipLaunchKernelGGL((KERNEL<16, Real, BL, 16, false>), grid, dim3(BL), 0, s, __VA_ARGS__);      \
    else if (NS == 20 && blocks)                                                                     \
      hipLaunchKernelGGL((KERNEL<32, Real, BL, 20, true>), grid, dim3(BL), 0, s, __VA_ARGS__);       \
    else if (NS == 24 && blocks)                                                                     \
      hipLaunchKernelGGL((KERNEL<32, Real, BL, 24, true>), grid, dim3(BL), 0, s, __VA_ARGS__);       \
    else if (NS == 20)                                                                               \
      hipLaunchKernelGGL((KERNEL<32, Real, BL, 20, false>), grid, dim3(BL), 0, s, __VA_ARGS__);      \
    else if (NS == 24)                                                                               \
      hipLaunchKernelGGL((KERNEL<32, Real, BL, 24, false>), grid, dim3(BL), 0, s, __VA_ARGS__);      \
    else                                                                                             \
      hipLaunchKernelGGL((KERNEL<32, Real, BL, 32, false>), grid, dim3(BL), 0, s, __VA_ARGS__);      \
  } while (0)

template <class Real>
hipError_t step_impl(const ModelTable<Real> *tab, int nj, int nr, int64_t B, int mode, float dt, float *q, float *qd,
                     const float *tau, int64_t tau_stride, float *out, float *save, int64_t save_stride,
                     hipStream_t s) {
  SAPAMD_TREE_DYN_LAUNCH(k_tree_step, tab, B, mode, dt, q, qd, tau, tau_stride, out, save, save_stride);
  return hipGetLastError();
}

template <class Real>
hipError_t step_vjp_impl(const ModelTable<Real> *tab, int nj, int nr, int64_t B, float dt, const float *state,
                         int64_t state_stride, const float *tau, int64_t tau_stride, float *aq, float *aqd,
                         float *gtau, hipStream_t s) {
  SAPAMD_TREE_DYN_LAUNCH(k_tree_step_vjp, tab, B, dt, state, state_stride, tau, tau_stride, aq, aqd, gtau);
  return hipGetLastError();
}

template <class Real>
hipError_t rollout_impl(const ModelTable<Real> *tab, int nj, int nr, int64_t B, int T, int K, float dt, float *q, float *qd,
                        const float *tau, float *ckpt, hipStream_t s) {
  SAPAMD_TREE_DYN_LAUNCH(k_tree_rollout, tab, B, T, K, dt, q, qd, tau, ckpt);
  return hipGetLastError();
}

template <class Real>
hipError_t rollout_vjp_impl(const ModelTable<Real> *tab, int nj, int nr, int64_t B, int T, int K, float dt, const float *tau,
                            const float *ckpt, float *seg, float *aq, float *aqd, float *gtau, hipStream_t s) {
  SAPAMD_TREE_DYN_LAUNCH(k_tree_rollout_vjp, tab, B, T, K, dt, tau, ckpt, seg, aq, aqd, gtau);
  return hipGetLastError();
}

}  // namespace

int tree_dyn_group(int nj) { return nj >= 1 && nj <= 8 ? 8 : nj <= 16 ? 16 : nj <= 32 ? 32 : 0; }

int tree_dyn_steps(int rows, int G) {
  if (G == 8) return 8;
  if (G == 16) return rows <= 12 ? 12 : 16;
  return rows <= 20 ? 20 : rows <= 24 ? 24 : 32;
}

DynFactorPlan dyn_factor_plan(const int *parent, int nj) {
  const int G = tree_dyn_group(nj);
  DynFactorPlan p{false, nj};
  if (!G) return p;
  // root subtrees: [j, j + size) in depth-first order for every joint j hanging from the universe
  int roots = 0, maxo = 0, maxsize = 0;
  for (int j = 0; j < nj; ++j) {
    if (parent[j] >= 0) continue;
    int e = j + 1;
    while (e < nj && parent[e] >= j) ++e;  // descendants follow their root contiguously
    ++roots;
    maxo = j;
    maxsize = e - j > maxsize ? e - j : maxsize;
  }
  const int steps = tree_dyn_steps(maxsize, G);
  if (G == 32 && roots > 1 && steps < tree_dyn_steps(nj, G) && maxo + steps <= G) p = {true, maxsize};
  return p;
}

hipError_t launch_tree_step(const ModelTable<float> *tab, const ModelTable<double> *tab64, int nj, int64_t B,
                            int mode, float dt, float *q, float *qd, const float *tau, int64_t tau_stride, float *out,
                            float *save, int64_t save_stride, hipStream_t s, int rows) {
  if (B <= 0) return hipSuccess;
  if (!tree_dyn_group(nj) || rows < 0 || rows > nj) return hipErrorNotSupported;
  const int nr = rows ? rows : nj;
  return tab64 ? step_impl<double>(tab64, nj, nr, B, mode, dt, q, qd, tau, tau_stride, out, save, save_stride, s)
               : step_impl<float>(tab, nj, nr, B, mode, dt, q, qd, tau, tau_stride, out, save, save_stride, s);
}

hipError_t launch_tree_step_vjp(const ModelTable<float> *tab, const ModelTable<double> *tab64, int nj, int64_t B,
                                float dt, const float *state, int64_t state_stride, const float *tau,
                                int64_t tau_stride, float *aq, float *aqd, float *gtau, hipStream_t s, int rows) {
  if (B <= 0) return hipSuccess;
  if (!tree_dyn_group(nj) || rows < 0 || rows > nj) return hipErrorNotSupported;
  const int nr = rows ? rows : nj;
  return tab64 ? step_vjp_impl<double>(tab64, nj, nr, B, dt, state, state_stride, tau, tau_stride, aq, aqd, gtau, s)
               : step_vjp_impl<float>(tab, nj, nr, B, dt, state, state_stride, tau, tau_stride, aq, aqd, gtau, s);
}

hipError_t launch_tree_rollout(const ModelTable<float> *tab, const ModelTable<double> *tab64, int nj, int64_t B,
                               int T, int K, float dt, float *q, float *qd, const float *tau, float *ckpt,
                               hipStream_t s, int rows) {
  if (B <= 0 || T <= 0) return hipSuccess;
  if (!tree_dyn_group(nj) || K < 1 || rows < 0 || rows > nj) return hipErrorNotSupported;
  const int nr = rows ? rows : nj;
  return tab64 ? rollout_impl<double>(tab64, nj, nr, B, T, K, dt, q, qd, tau, ckpt, s)
               : rollout_impl<float>(tab, nj, nr, B, T, K, dt, q, qd, tau, ckpt, s);
}

hipError_t launch_tree_rollout_vjp(const ModelTable<float> *tab, const ModelTable<double> *tab64, int nj, int64_t B,
                                   int T, int K, float dt, const float *tau, const float *ckpt, float *seg, float *aq,
                                   float *aqd, float *gtau, hipStream_t s, int rows) {
  if (B <= 0 || T <= 0) return hipSuccess;
  if (!tree_dyn_group(nj) || K < 1 || rows < 0 || rows > nj) return hipErrorNotSupported;
  const int nr = rows ? rows : nj;
  return tab64 ? rollout_vjp_impl<double>(tab64, nj, nr, B, T, K, dt, tau, ckpt, seg, aq, aqd, gtau, s)
               : rollout_vjp_impl<float>(tab, nj, nr, B, T, K, dt, tau, ckpt, seg, aq, aqd, gtau, s);
}

hipError_t launch_rollout_mixed(RolloutMixArgs A, bool vjp, hipStream_t s) {
  if (A.nseg < 1 || A.nseg > kMaxSegments || A.K < 1 || A.T < 0) return hipErrorInvalidValue;
  if (A.T == 0) return hipSuccess;
  // heaviest first: FP64 before FP32, then by lanes per environment and joints
  RolloutSegment *b = A.seg, *e = A.seg + A.nseg;
  for (RolloutSegment *x = b; x != e; ++x) {
    if (!tree_dyn_group(x->nj) || x->B < 0 || x->rows < 0 || x->rows > x->nj) return hipErrorInvalidValue;
    if (x->rows && x->rows < x->nj &&
        (tree_dyn_group(x->nj) != 32 || tree_dyn_steps(x->rows, 32) > 24))
      return hipErrorInvalidValue;
    x->variant = mixed_variant(x->nj, x->rows ? x->rows : x->nj, x->fp64 != 0);
  }
  std::stable_sort(b, e, [](const RolloutSegment &x, const RolloutSegment &y) {
    return x.variant != y.variant ? x.variant > y.variant : x.nj > y.nj;
  });
  int64_t blocks = 0;
  size_t lds = 0;
  int n = 0;
  for (int i = 0; i < A.nseg; ++i) {
    if (A.seg[i].B == 0) continue;
    RolloutSegment &x = A.seg[n++] = A.seg[i];
    const int G = tree_dyn_group(x.nj);
    x.block0 = blocks;
    blocks += (x.B * G + kMixBlock - 1) / kMixBlock;
    lds = std::max(lds, mixed_lds_bytes(x.variant));
  }
  A.nseg = n;
  if (blocks == 0) return hipSuccess;
  if (blocks > INT32_MAX) return hipErrorInvalidValue;
  if (vjp)
    hipLaunchKernelGGL(k_tree_rollout_mixed<true>, dim3((unsigned)blocks), dim3(kMixBlock), (unsigned)lds, s, A);
  else
    hipLaunchKernelGGL(k_tree_rollout_mixed<false>, dim3((unsigned)blocks), dim3(kMixBlock), (unsigned)lds, s, A);
  return hipGetLastError();
}

}  // namespace sapamd
