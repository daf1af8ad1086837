// Lanes-per-joint forward dynamics and rollout steps for trees of <= 32
// joints (FP32): one environment per group of G lanes, lane j owning joint j,
// everything in registers and LDS (no private segment).
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

#include "lanes.h"
#include "launch.h"
#include "spatial.h"
#include "tree.h"

namespace sapamd {

namespace {

using namespace lanes;

constexpr int kTreeDynJoints = 32;
constexpr int kDynTableWords4 = (int)((kHeaderBytes + sizeof(JointRec<float>) * kTreeDynJoints) / 16);

// Orders one wavefront's LDS accesses across lanes: the fences keep the
// compiler from moving loads or stores across the point (a wavefront's own LDS
// operations complete in order, so no wait is emitted at wavefront scope).
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// LDL^T of the symmetric positive definite n x n matrix in A (lower triangle
// used), lane i owning row i: after the call A holds L below the diagonal and
// D on it.
template <int G> __device__ __forceinline__ void ldlt_factor(float (*A)[G + 1], int i, int n) {
  for (int k = 0; k < n; ++k) {
    wave_sync();
    const float dk = A[k][k];
    const bool act = i > k && i < n;
    const float l = act ? A[i][k] / dk : 0.f;
    if (act)
      for (int j = k + 1; j <= i; ++j) A[i][j] -= l * A[j][k];
    wave_sync();
    if (act) A[i][k] = l;
  }
  wave_sync();
}
// x = (L D L^T)^-1 r with lane i holding r_i; returns x_i. y: G floats of LDS.
template <int G> __device__ __forceinline__ float ldlt_solve(float (*A)[G + 1], float *y, int i, int n, float r) {
  for (int k = 0; k < n; ++k) {
    if (i == k) y[k] = r;
    wave_sync();
    const float yk = y[k];
    if (i > k && i < n) r -= A[i][k] * yk;
    wave_sync();
  }
  float z = i < n ? r / A[i][i] : 0.f;
  for (int k = n - 1; k >= 0; --k) {
    if (i == k) y[k] = z;
    wave_sync();
    const float xk = y[k];
    if (i < k) z -= A[k][i] * xk;
    wave_sync();
  }
  return z;
}

// Rows padded by one word: lane i walking row i and lane j walking row j hit
// different banks.
template <int G> struct DynLds {
  float A[G][G + 1];
  float y[G];
};

// One environment's kinematics and M / bias; the factorization is left in L.
template <int G> struct TreeFd {
  Tree<G> T;
  int g, nj;
  bool valid;
  const JointRec<float> *J;
  Xf X;
  S6<float> col, v;

  // q, qd in joint order (this lane's joint); writes M (row g) and returns the
  // bias b_g = RNEA(q, qd, 0)_g. The caller factors L->A.
  __device__ __forceinline__ float setup(const ModelTable<float> *M, float qj, float qdj, DynLds<G> &L) {
    float s, c;
    joint_sincos(qj, &s, &c);
    joint_world(*J, qj, s, c, static_cast<const float *>(nullptr), V3<float>{0, 0, 0}, true, X.R, X.p);
    if (!valid) X = identity_xf();
    T.prefix(X);
    col = valid ? jcol(*J, X.R, X.p) : zero6<float>();
    v = qdj * col;
    T.prefix(v);
    S6<float> a = qdj * crossm(v, col);
    T.prefix(a);
    a.l = a.l - V3<float>{M->gravity[0], M->gravity[1], M->gravity[2]};
    const WInertia<float> Y = world_inertia_fast(*J, X.R, X.p);
    S6<float> f = valid ? inertia_apply(Y, a) + crossf(v, inertia_apply(Y, v)) : zero6<float>();
    T.subtree(f);
    const float bias = dot6(col, f);
    // composite inertia of the subtree and the lower triangle of row g of M
    AInertia<float> Ya = additive(Y);
    float y[10] = {Ya.m, Ya.h.x, Ya.h.y, Ya.h.z, Ya.I[0], Ya.I[1], Ya.I[2], Ya.I[3], Ya.I[4], Ya.I[5]};
    if (!valid)
#pragma unroll
      for (int k = 0; k < 10; ++k) y[k] = 0.f;
    T.template subtree<10>(y);
    AInertia<float> Yc;
    Yc.m = y[0];
    Yc.h = {y[1], y[2], y[3]};
#pragma unroll
    for (int k = 0; k < 6; ++k) Yc.I[k] = y[4 + k];
    const S6<float> Fg = apply(Yc, col);
    for (int i = 0; i < nj; ++i) {  // uniform loop: every lane takes part in the permutes
      const S6<float> ci = bperm6(T.base4 + 4 * i, col);
      const JointRec<float> &Ji = joints_of(M)[i];
      const bool anc = i <= g && g < i + Ji.subtree;  // i ancestor-or-self of g
      if (valid && i <= g) L.A[g][i] = anc ? dot6(ci, Fg) : 0.f;
    }
    return bias;
  }
};

template <int G>
__device__ __forceinline__ void init_fd(TreeFd<G> &F, const ModelTable<float> *M, int g) {
  F.g = g;
  F.nj = M->nj;
  F.valid = g < F.nj;
  F.J = joints_of(M) + (F.valid ? g : F.nj - 1);
  F.T.g = g;
  F.T.base4 = ((threadIdx.x & 63) & ~(G - 1)) * 4;
#pragma unroll
  for (int k = 0; k < Tree<G>::K; ++k) F.T.anc[k] = F.valid ? F.J->anc[k] : -1;
  F.T.size = F.valid ? F.J->subtree : 0;
}

// Stage the model table in LDS (block-wide) and return the LDS copy.
__device__ __forceinline__ const ModelTable<float> *stage_table(const ModelTable<float> *Mg, float4 *s_tab) {
  const int nj = Mg->nj;
  const int n4 = (int)((kHeaderBytes + sizeof(JointRec<float>) * nj) / 16);
  const float4 *src = reinterpret_cast<const float4 *>(Mg);
  for (int i = threadIdx.x; i < n4; i += blockDim.x) s_tab[i] = src[i];
  __syncthreads();
  return reinterpret_cast<const ModelTable<float> *>(s_tab);
}

// mode 0: forward dynamics qdd = FD(q, qd, tau) into out (B, nv)
// mode 1: rollout step in place on q, qd; save (pre-step q | qd) when given
template <int G>
__global__ __launch_bounds__(256) void k_tree_step(const ModelTable<float> *__restrict__ Mg, int64_t B, int mode,
                                                   float dt, float *__restrict__ q, float *__restrict__ qd,
                                                   const float *__restrict__ tau, int64_t tau_stride,
                                                   float *__restrict__ out, float *__restrict__ save,
                                                   int64_t save_stride) {
  __shared__ float4 s_tab[kDynTableWords4];
  __shared__ DynLds<G> s_dyn[256 / G];
  const int g = threadIdx.x & (G - 1);
  const int64_t env = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) / G;
  const bool live = env < B;
  const int64_t e = live ? env : B - 1;
  const int nv = Mg->nv;
  // loads first (one round trip): lane g reads user dof g
  const int gc = g < nv ? g : nv - 1;
  const float qu = q[e * nv + gc], qdu = qd[e * nv + gc], tu = tau[e * tau_stride + gc];
  const ModelTable<float> *M = stage_table(Mg, s_tab);
  DynLds<G> &L = s_dyn[threadIdx.x / G];
  TreeFd<G> F;
  init_fd<G>(F, M, g);
  const int src = F.T.base4 + 4 * F.J->vuser;
  float qj = bperm(src, qu), qdj = bperm(src, qdu), tj = bperm(src, tu);
  if (!F.valid) qj = qdj = tj = 0.f;
  const float bias = F.setup(M, qj, qdj, L);
  ldlt_factor<G>(L.A, g, F.nj);
  const float qddj = ldlt_solve<G>(L.A, L.y, g, F.nj, tj - bias);
  if (!live) return;
  if (mode == 0) {
    if (F.valid) out[e * nv + F.J->vuser] = qddj;
    return;
  }
  if (save && g < nv) {
    save[e * save_stride + g] = qu;
    save[e * save_stride + nv + g] = qdu;
  }
  if (F.valid) {
    const float v1 = qdj + dt * qddj;
    qd[e * nv + F.J->vuser] = v1;
    q[e * nv + F.J->vuser] = qj + dt * v1;
  }
}

// Adjoint of one step; state = pre-step (q | qd) rows, aq / aqd hold the
// post-step adjoint on entry and the pre-step adjoint on exit.
template <int G>
__global__ __launch_bounds__(256) void k_tree_step_vjp(const ModelTable<float> *__restrict__ Mg, int64_t B, float dt,
                                                       const float *__restrict__ state, int64_t state_stride,
                                                       const float *__restrict__ tau, int64_t tau_stride,
                                                       float *__restrict__ aq, float *__restrict__ aqd,
                                                       float *__restrict__ gtau) {
  __shared__ float4 s_tab[kDynTableWords4];
  __shared__ DynLds<G> s_dyn[256 / G];
  const int g = threadIdx.x & (G - 1);
  const int64_t env = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) / G;
  const bool live = env < B;
  const int64_t e = live ? env : B - 1;
  const int nv = Mg->nv;
  const int gc = g < nv ? g : nv - 1;
  const float qu = state[e * state_stride + gc], qdu = state[e * state_stride + nv + gc];
  const float tu = tau[e * tau_stride + gc], aqu = aq[e * nv + gc], aqdu = aqd[e * nv + gc];
  const ModelTable<float> *M = stage_table(Mg, s_tab);
  DynLds<G> &L = s_dyn[threadIdx.x / G];
  TreeFd<G> F;
  init_fd<G>(F, M, g);
  const int src = F.T.base4 + 4 * F.J->vuser;
  float qj = bperm(src, qu), qdj = bperm(src, qdu), tj = bperm(src, tu);
  float aqj = bperm(src, aqu), aqdj = bperm(src, aqdu);
  if (!F.valid) qj = qdj = tj = aqj = aqdj = 0.f;
  const float bias = F.setup(M, qj, qdj, L);
  ldlt_factor<G>(L.A, g, F.nj);
  const float qddj = ldlt_solve<G>(L.A, L.y, g, F.nj, tj - bias);
  const float cj = aqdj + dt * aqj;
  const float lam = ldlt_solve<G>(L.A, L.y, g, F.nj, dt * cj);

  // RNEA vector-Jacobian product at (q, qd, qdd) for the upstream lambda
  const Tree<G> &T = F.T;
  const S6<float> col = F.col, v = F.v;
  S6<float> a = qddj * col + qdj * crossm(v, col);
  T.prefix(a);
  a.l = a.l - V3<float>{M->gravity[0], M->gravity[1], M->gravity[2]};
  const WInertia<float> Y = world_inertia_fast(*F.J, F.X.R, F.X.p);
  S6<float> Ff = F.valid ? inertia_apply(Y, a) + crossf(v, inertia_apply(Y, v)) : zero6<float>();
  T.subtree(Ff);
  const float gj = F.valid ? lam : 0.f;
  S6<float> w = gj * col;
  T.prefix(w);
  const S6<float> Iw = inertia_apply(Y, w), Iv = inertia_apply(Y, v);
  S6<float> Hw = zero6<float>(), QS = zero6<float>();
  if (F.valid) {
    Hw = Iw;
    QS = crossf(w, Iv) - crossf(v, Iw) + inertia_apply(Y, crossm(v, w));
  }
  float acc[12] = {Hw.l.x, Hw.l.y, Hw.l.z, Hw.a.x, Hw.a.y, Hw.a.z, QS.l.x, QS.l.y, QS.l.z, QS.a.x, QS.a.y, QS.a.z};
  T.template subtree<12>(acc);
  Hw = {{acc[0], acc[1], acc[2]}, {acc[3], acc[4], acc[5]}};
  QS = {{acc[6], acc[7], acc[8]}, {acc[9], acc[10], acc[11]}};
  if (!live || !F.valid) return;
  const S6<float> wl = w - gj * col, vl = v - qdj * col;
  const S6<float> al = a - qddj * col - qdj * crossm(v, col);
  const float dq = dot6(col, crossf(vl, QS + crossf(vl, Hw)) - crossf(wl, Ff) - crossf(al, Hw));
  const float dqd = -dot6(col, QS + 2.f * crossf(vl, Hw));
  const int u = F.J->vuser;
  aq[e * nv + u] = aqj - dq;
  aqd[e * nv + u] = cj - dqd;
  gtau[e * tau_stride + u] = lam;
}

}  // namespace

int tree_dyn_group(int nj) { return nj >= 1 && nj <= 8 ? 8 : nj <= 16 ? 16 : nj <= 32 ? 32 : 0; }

hipError_t launch_tree_step(const ModelTable<float> *tab, int nj, int64_t B, int mode, float dt, float *q, float *qd,
                            const float *tau, int64_t tau_stride, float *out, float *save, int64_t save_stride,
                            hipStream_t s) {
  if (B <= 0) return hipSuccess;
  const int G = tree_dyn_group(nj);
  if (!G) return hipErrorNotSupported;
  const dim3 grid((unsigned)((B * G + 255) / 256));
  if (G == 8)
    hipLaunchKernelGGL(k_tree_step<8>, grid, dim3(256), 0, s, tab, B, mode, dt, q, qd, tau, tau_stride, out, save,
                       save_stride);
  else if (G == 16)
    hipLaunchKernelGGL(k_tree_step<16>, grid, dim3(256), 0, s, tab, B, mode, dt, q, qd, tau, tau_stride, out, save,
                       save_stride);
  else
    hipLaunchKernelGGL(k_tree_step<32>, grid, dim3(256), 0, s, tab, B, mode, dt, q, qd, tau, tau_stride, out, save,
                       save_stride);
  return hipGetLastError();
}

hipError_t launch_tree_step_vjp(const ModelTable<float> *tab, int nj, int64_t B, float dt, const float *state,
                                int64_t state_stride, const float *tau, int64_t tau_stride, float *aq, float *aqd,
                                float *gtau, hipStream_t s) {
  if (B <= 0) return hipSuccess;
  const int G = tree_dyn_group(nj);
  if (!G) return hipErrorNotSupported;
  const dim3 grid((unsigned)((B * G + 255) / 256));
  if (G == 8)
    hipLaunchKernelGGL(k_tree_step_vjp<8>, grid, dim3(256), 0, s, tab, B, dt, state, state_stride, tau, tau_stride, aq,
                       aqd, gtau);
  else if (G == 16)
    hipLaunchKernelGGL(k_tree_step_vjp<16>, grid, dim3(256), 0, s, tab, B, dt, state, state_stride, tau, tau_stride,
                       aq, aqd, gtau);
  else
    hipLaunchKernelGGL(k_tree_step_vjp<32>, grid, dim3(256), 0, s, tab, B, dt, state, state_stride, tau, tau_stride,
                       aq, aqd, gtau);
  return hipGetLastError();
}

}  // namespace sapamd
