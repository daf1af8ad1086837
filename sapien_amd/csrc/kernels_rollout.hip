// Semi-implicit Euler rollouts with ABA forward dynamics and their adjoint
// (thread-per-environment, any tree up to 64 joints; FP32 ABA, or FP64
// factorized forward dynamics for ill-conditioned trees).
//
// Step t:  qdd_t = ABA(q_t, qd_t, tau_t);  qd_{t+1} = qd_t + dt qdd_t;  q_{t+1} = q_t + dt qd_{t+1}.
// Adjoint of a step, from RNEA(q, qd, qdd) = tau (implicit differentiation):
//   c       = abar_qd' + dt abar_q'
//   lambda  = M(q_t)^-1 (dt c)            (ABA with qd = 0, gravity off, tau = dt c)
//   dL/dtau_t = lambda
//   dL/dq_t   = abar_q' - (dRNEA/dq)^T lambda
//   dL/dqd_t  = c       - (dRNEA/dqd)^T lambda
// with the RNEA vector-Jacobian product of kernels_adjoint.hip evaluated at
// (q_t, qd_t, qdd_t). Each launch advances (or back-propagates) one step, so a
// rollout is a sequence of small launches that a caller captures in a HIP graph.
// The checkpointing schedule (states every K steps, segments recomputed in the
// backward pass) lives in the host driver.
#include <hip/hip_runtime.h>

#include "launch.h"
#include "spatial.h"

namespace sapamd {

namespace {

// World poses and Jacobian columns of every joint for configuration q (user order).
template <class T, int MJ>
__device__ __forceinline__ void fk_cols(const ModelTable<T> *M, const T *q, T (&oR)[MJ][9], V3<T> (&op)[MJ],
                                        S6<T> (&col)[MJ]) {
  const JointRec<T> *JR = joints_of(M);
  for (int j = 0; j < M->nj; ++j) {
    const JointRec<T> &J = JR[j];
    const int p = J.parent;
    const T qj = q[J.vuser];
    T s, c;
    sincos_t(qj, &s, &c);
    joint_world(J, qj, s, c, p >= 0 ? oR[p] : static_cast<const T *>(nullptr), p >= 0 ? op[p] : V3<T>{0, 0, 0},
                p < 0, oR[j], op[j]);
    col[j] = jcol(J, oR[j], op[j]);
  }
}

// Articulated-body algorithm in the world frame (three passes, no frame
// changes: articulated inertias of children add directly into the parent).
// qd may be nullptr (zero); gravity selects a_0 = -g or 0. Returns qdd (user order).
template <class T, int MJ>
__device__ __forceinline__ void aba_world(const ModelTable<T> *M, const T (&oR)[MJ][9], const V3<T> (&op)[MJ],
                                          const S6<T> (&col)[MJ], const T *qd, const T *tau, bool gravity, T *qdd) {
  const JointRec<T> *JR = joints_of(M);
  const int nj = M->nj;
  S6<T> vel[MJ], cb[MJ], pA[MJ], U[MJ];
  T IA[MJ][21];  // symmetric 6x6, upper triangle row-major
  T D[MJ], uu[MJ];
  for (int j = 0; j < nj; ++j) {
    const JointRec<T> &J = JR[j];
    const int p = J.parent;
    const T qdj = qd ? qd[J.vuser] : T(0);
    vel[j] = (p >= 0 ? vel[p] : zero6<T>()) + qdj * col[j];
    cb[j] = qdj * crossm(vel[j], col[j]);
    const WInertia<T> Y = world_inertia(J, oR[j], op[j]);
    pA[j] = crossf(vel[j], inertia_apply(Y, vel[j]));
    const AInertia<T> A = additive(Y);
    // [[m I, -[h]x], [[h]x, I_O]] upper triangle
    T *I = IA[j];
    const T m = A.m;
    I[0] = m;  I[1] = 0;  I[2] = 0;  I[3] = 0;       I[4] = A.h.z;  I[5] = -A.h.y;
    I[6] = m;  I[7] = 0;  I[8] = -A.h.z; I[9] = 0;   I[10] = A.h.x;
    I[11] = m; I[12] = A.h.y; I[13] = -A.h.x; I[14] = 0;
    I[15] = A.I[0]; I[16] = A.I[1]; I[17] = A.I[2];
    I[18] = A.I[3]; I[19] = A.I[4];
    I[20] = A.I[5];
  }
  auto idx = [](int r, int c) { return r <= c ? r * 6 - r * (r - 1) / 2 + (c - r) : c * 6 - c * (c - 1) / 2 + (r - c); };
  for (int j = nj - 1; j >= 0; --j) {
    const T *I = IA[j];
    T s[6] = {col[j].l.x, col[j].l.y, col[j].l.z, col[j].a.x, col[j].a.y, col[j].a.z}, Uv[6];
    for (int r = 0; r < 6; ++r) {
      T acc = T(0);
      for (int k = 0; k < 6; ++k) acc += I[idx(r, k)] * s[k];
      Uv[r] = acc;
    }
    U[j] = {{Uv[0], Uv[1], Uv[2]}, {Uv[3], Uv[4], Uv[5]}};
    D[j] = dot6(col[j], U[j]);
    uu[j] = tau[JR[j].vuser] - dot6(col[j], pA[j]);
    const int p = JR[j].parent;
    if (p < 0) continue;
    const T invD = T(1) / D[j];
    // Ia = IA - U U^T / D;  pa = pA + Ia c + U u / D
    T Ia[21];
    for (int r = 0; r < 6; ++r)
      for (int c = r; c < 6; ++c) Ia[idx(r, c)] = I[idx(r, c)] - Uv[r] * Uv[c] * invD;
    T cv[6] = {cb[j].l.x, cb[j].l.y, cb[j].l.z, cb[j].a.x, cb[j].a.y, cb[j].a.z}, pv[6];
    for (int r = 0; r < 6; ++r) {
      T acc = T(0);
      for (int k = 0; k < 6; ++k) acc += Ia[idx(r, k)] * cv[k];
      pv[r] = acc + Uv[r] * uu[j] * invD;
    }
    for (int k = 0; k < 21; ++k) IA[p][k] += Ia[k];
    pA[p] = pA[p] + pA[j] + S6<T>{{pv[0], pv[1], pv[2]}, {pv[3], pv[4], pv[5]}};
  }
  const S6<T> a0 = gravity ? S6<T>{{-M->gravity[0], -M->gravity[1], -M->gravity[2]}, {T(0), T(0), T(0)}} : zero6<T>();
  S6<T> acc[MJ];
  for (int j = 0; j < nj; ++j) {
    const int p = JR[j].parent;
    acc[j] = (p >= 0 ? acc[p] : a0) + cb[j];
    const T x = (uu[j] - dot6(U[j], acc[j])) / D[j];
    qdd[JR[j].vuser] = x;
    acc[j] = acc[j] + x * col[j];
  }
}

// RNEA vector-Jacobian product at (q, qd, qdd) for the upstream gradient g:
// dq += -(dtau/dq)^T g, dqd += -(dtau/dqd)^T g (signs folded for the rollout).
template <class T, int MJ>
__device__ __forceinline__ void rnea_vjp_sub(const ModelTable<T> *M, const T (&oR)[MJ][9], const V3<T> (&op)[MJ],
                                             const S6<T> (&col)[MJ], const T *qd, const T *qdd, const T *g,
                                             T *dq_acc, T *dqd_acc) {
  const JointRec<T> *JR = joints_of(M);
  const int nj = M->nj;
  S6<T> vel[MJ], acc[MJ], F[MJ], W[MJ], Hw[MJ], QS[MJ];
  const S6<T> a0 = {{-M->gravity[0], -M->gravity[1], -M->gravity[2]}, {T(0), T(0), T(0)}};
  for (int j = 0; j < nj; ++j) {
    const JointRec<T> &J = JR[j];
    const int p = J.parent, u = J.vuser;
    vel[j] = (p >= 0 ? vel[p] : zero6<T>()) + qd[u] * col[j];
    acc[j] = (p >= 0 ? acc[p] : a0) + qdd[u] * col[j] + qd[u] * crossm(vel[j], col[j]);
    W[j] = (p >= 0 ? W[p] : zero6<T>()) + g[u] * col[j];
    const WInertia<T> Y = world_inertia(J, oR[j], op[j]);
    const S6<T> Iv = inertia_apply(Y, vel[j]), Iw = inertia_apply(Y, W[j]);
    F[j] = inertia_apply(Y, acc[j]) + crossf(vel[j], Iv);
    Hw[j] = Iw;
    QS[j] = crossf(W[j], Iv) - crossf(vel[j], Iw) + inertia_apply(Y, crossm(vel[j], W[j]));
  }
  for (int j = nj - 1; j >= 0; --j) {
    const int p = JR[j].parent;
    if (p < 0) continue;
    F[p] = F[p] + F[j];
    Hw[p] = Hw[p] + Hw[j];
    QS[p] = QS[p] + QS[j];
  }
  for (int k = 0; k < nj; ++k) {
    const int u = JR[k].vuser;
    const S6<T> c = col[k];
    const S6<T> wl = W[k] - g[u] * c, vl = vel[k] - qd[u] * c;
    const S6<T> al = acc[k] - qdd[u] * c - qd[u] * crossm(vel[k], c);
    dq_acc[u] -= dot6(c, crossf(vl, QS[k] + crossf(vl, Hw[k])) - crossf(wl, F[k]) - crossf(al, Hw[k]));
    dqd_acc[u] -= -dot6(c, QS[k] + T(2) * crossf(vl, Hw[k]));
  }
}

// ---------------------------------------------------------------------------
// FP64 forward dynamics by factorization, for trees whose mass matrix is
// ill-conditioned (cond(M) > 1e6, where FP32 ABA loses too many digits):
// M(q) by the composite-rigid-body algorithm, the bias b(q, qd) = RNEA(q, qd, 0)
// (gravity included), then Featherstone's LTDL factorization M = L^T D L,
// which keeps the tree's sparsity (row k of L is nonzero only at ancestors of
// k, so there is no fill-in), and the matching back-substitution.
// Everything in FP64; joints in internal (depth-first) order, parents first.
template <int MJ> struct Fd64 {
  double H[MJ][MJ];  // H[k][i] for i ancestor-or-self of k; LTDL factors in place
  int par[MJ];
  int nj;

  __device__ __forceinline__ void factor() {
    for (int k = nj - 1; k >= 0; --k)
      for (int i = par[k]; i >= 0; i = par[i]) {
        const double a = H[k][i] / H[k][k];
        for (int j = i; j >= 0; j = par[j]) H[i][j] -= a * H[k][j];
        H[k][i] = a;
      }
  }
  // y <- M^-1 y (internal order)
  __device__ __forceinline__ void solve(double *y) const {
    for (int i = nj - 1; i >= 0; --i)
      for (int j = par[i]; j >= 0; j = par[j]) y[j] -= H[i][j] * y[i];
    for (int i = 0; i < nj; ++i) y[i] /= H[i][i];
    for (int i = 0; i < nj; ++i)
      for (int j = par[i]; j >= 0; j = par[j]) y[i] -= H[i][j] * y[j];
  }
};

// Builds M (into F.H) and b = RNEA(q, qd, 0) (internal order) for user-order q, qd.
template <int MJ>
__device__ __forceinline__ void mass_and_bias_f64(const ModelTable<double> *M, const double *q, const double *qd,
                                                  Fd64<MJ> &F, double *bias) {
  const JointRec<double> *JR = joints_of(M);
  const int nj = M->nj;
  F.nj = nj;
  double oR[MJ][9];
  V3<double> op[MJ];
  S6<double> col[MJ], vel[MJ], acc[MJ], f[MJ];
  AInertia<double> Yc[MJ];
  const S6<double> a0 = {{-M->gravity[0], -M->gravity[1], -M->gravity[2]}, {0.0, 0.0, 0.0}};
  for (int j = 0; j < nj; ++j) {
    const JointRec<double> &J = JR[j];
    const int p = J.parent, u = J.vuser;
    F.par[j] = p;
    double s, c;
    sincos_t(q[u], &s, &c);
    joint_world(J, q[u], s, c, p >= 0 ? oR[p] : static_cast<const double *>(nullptr),
                p >= 0 ? op[p] : V3<double>{0, 0, 0}, p < 0, oR[j], op[j]);
    col[j] = jcol(J, oR[j], op[j]);
    const WInertia<double> Y = world_inertia(J, oR[j], op[j]);
    vel[j] = (p >= 0 ? vel[p] : zero6<double>()) + qd[u] * col[j];
    acc[j] = (p >= 0 ? acc[p] : a0) + qd[u] * crossm(vel[j], col[j]);
    f[j] = inertia_apply(Y, acc[j]) + crossf(vel[j], inertia_apply(Y, vel[j]));
    Yc[j] = additive(Y);
  }
  for (int j = nj - 1; j > 0; --j) {
    const int p = JR[j].parent;
    if (p < 0) continue;
    add_to(Yc[p], Yc[j]);
    f[p] = f[p] + f[j];
  }
  for (int j = 0; j < nj; ++j) {
    bias[j] = dot6(col[j], f[j]);
    const S6<double> Fj = apply(Yc[j], col[j]);
    for (int i = j; i >= 0; i = JR[i].parent) F.H[j][i] = dot6(col[i], Fj);
  }
}

// qdd = M^-1 (tau - b) and, when rhs2 != nullptr, out2 = M^-1 rhs2 (user order,
// FP32 in and out, FP64 inside).
template <class T, int MJ>
__device__ __forceinline__ void fd_ldlt(const ModelTable<double> *M, const T *q, const T *qd, const T *tau,
                                        const T *rhs2, T *qdd, T *out2) {
  const JointRec<double> *JR = joints_of(M);
  const int nj = M->nj;
  double qq[MJ], qv[MJ], y[MJ], b[MJ];
  for (int i = 0; i < nj; ++i) {
    qq[i] = double(q[i]);
    qv[i] = double(qd[i]);
  }
  Fd64<MJ> F;
  mass_and_bias_f64<MJ>(M, qq, qv, F, b);
  F.factor();
  for (int j = 0; j < nj; ++j) y[j] = double(tau[JR[j].vuser]) - b[j];
  F.solve(y);
  for (int j = 0; j < nj; ++j) qdd[JR[j].vuser] = T(y[j]);
  if (rhs2) {
    for (int j = 0; j < nj; ++j) y[j] = double(rhs2[JR[j].vuser]);
    F.solve(y);
    for (int j = 0; j < nj; ++j) out2[JR[j].vuser] = T(y[j]);
  }
}

template <class T, int MJ, bool F64>
__global__ __launch_bounds__(64) void k_rollout_step(const ModelTable<T> *__restrict__ M,
                                                     const ModelTable<double> *__restrict__ M64, int64_t B, T dt,
                                                     T *__restrict__ q, T *__restrict__ qd,
                                                     const T *__restrict__ tau, int64_t tau_stride,
                                                     T *__restrict__ save, int64_t save_stride) {
  const int64_t b = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (b >= B) return;
  const int nv = M->nv;
  T qb[MJ], qdb[MJ], tb[MJ], qdd[MJ];
  for (int i = 0; i < nv; ++i) {
    qb[i] = q[b * nv + i];
    qdb[i] = qd[b * nv + i];
    tb[i] = tau[b * tau_stride + i];
  }
  if (save) {
    for (int i = 0; i < nv; ++i) {
      save[b * save_stride + i] = qb[i];
      save[b * save_stride + nv + i] = qdb[i];
    }
  }
  if constexpr (F64) {
    fd_ldlt<T, MJ>(M64, qb, qdb, tb, static_cast<const T *>(nullptr), qdd, static_cast<T *>(nullptr));
  } else {
    T oR[MJ][9];
    V3<T> op[MJ];
    S6<T> col[MJ];
    fk_cols<T, MJ>(M, qb, oR, op, col);
    aba_world<T, MJ>(M, oR, op, col, qdb, tb, true, qdd);
  }
  for (int i = 0; i < nv; ++i) {
    const T v = qdb[i] + dt * qdd[i];
    qd[b * nv + i] = v;
    q[b * nv + i] = qb[i] + dt * v;
  }
}

template <class T, int MJ, bool F64>
__global__ __launch_bounds__(64) void k_rollout_step_vjp(const ModelTable<T> *__restrict__ M,
                                                         const ModelTable<double> *__restrict__ M64, int64_t B, T dt,
                                                         const T *__restrict__ state, int64_t state_stride,
                                                         const T *__restrict__ tau, int64_t tau_stride,
                                                         T *__restrict__ aq, T *__restrict__ aqd,
                                                         T *__restrict__ gtau) {
  const int64_t b = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (b >= B) return;
  const int nv = M->nv;
  T qb[MJ], qdb[MJ], tb[MJ], qdd[MJ], rhs[MJ], lam[MJ], dq[MJ], dqd[MJ];
  for (int i = 0; i < nv; ++i) {
    qb[i] = state[b * state_stride + i];
    qdb[i] = state[b * state_stride + nv + i];
    tb[i] = tau[b * tau_stride + i];
    const T c = aqd[b * nv + i] + dt * aq[b * nv + i];
    dqd[i] = c;
    dq[i] = aq[b * nv + i];
    rhs[i] = dt * c;
  }
  T oR[MJ][9];
  V3<T> op[MJ];
  S6<T> col[MJ];
  fk_cols<T, MJ>(M, qb, oR, op, col);
  if constexpr (F64) {
    fd_ldlt<T, MJ>(M64, qb, qdb, tb, rhs, qdd, lam);
  } else {
    aba_world<T, MJ>(M, oR, op, col, qdb, tb, true, qdd);
    aba_world<T, MJ>(M, oR, op, col, static_cast<const T *>(nullptr), rhs, false, lam);
  }
  rnea_vjp_sub<T, MJ>(M, oR, op, col, qdb, qdd, lam, dq, dqd);
  for (int i = 0; i < nv; ++i) {
    aq[b * nv + i] = dq[i];
    aqd[b * nv + i] = dqd[i];
    gtau[b * tau_stride + i] = lam[i];
  }
}

// Batched forward dynamics through the FP64 factorization (FP32 in and out).
template <int MJ>
__global__ __launch_bounds__(64) void k_fd_ldlt(const ModelTable<double> *__restrict__ M64, int64_t B,
                                                const float *__restrict__ q, const float *__restrict__ qd,
                                                const float *__restrict__ tau, float *__restrict__ qdd) {
  const int64_t b = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (b >= B) return;
  const int nv = M64->nv;
  float qb[MJ], qdb[MJ], tb[MJ], out[MJ];
  for (int i = 0; i < nv; ++i) {
    qb[i] = q[b * nv + i];
    qdb[i] = qd[b * nv + i];
    tb[i] = tau[b * nv + i];
  }
  fd_ldlt<float, MJ>(M64, qb, qdb, tb, static_cast<const float *>(nullptr), out, static_cast<float *>(nullptr));
  for (int i = 0; i < nv; ++i) qdd[b * nv + i] = out[i];
}

inline int pick(int nj) { return nj <= 8 ? 8 : nj <= 16 ? 16 : nj <= 24 ? 24 : nj <= 32 ? 32 : nj <= 48 ? 48 : 64; }

__global__ void k_restore(int64_t B, int nv, const float *__restrict__ src, int64_t stride, float *__restrict__ q,
                          float *__restrict__ qd) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= B * nv) return;
  const int64_t b = i / nv, k = i % nv;
  q[i] = src[b * stride + k];
  qd[i] = src[b * stride + nv + k];
}

}  // namespace

hipError_t launch_restore_state(int64_t B, int nv, const float *src, int64_t stride, float *q, float *qd,
                                hipStream_t s) {
  if (B <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_restore, dim3((unsigned)((B * nv + 255) / 256)), dim3(256), 0, s, B, nv, src, stride, q, qd);
  return hipGetLastError();
}

#define SAPAMD_ROLLOUT_DISPATCH(KERNEL, ...)                                                          \
  switch (pick(nj) * 2 + (tab64 ? 1 : 0)) {                                                            \
    case 16: hipLaunchKernelGGL((KERNEL<T, 8, false>), grid, dim3(64), 0, s, __VA_ARGS__); break;      \
    case 17: hipLaunchKernelGGL((KERNEL<T, 8, true>), grid, dim3(64), 0, s, __VA_ARGS__); break;       \
    case 32: hipLaunchKernelGGL((KERNEL<T, 16, false>), grid, dim3(64), 0, s, __VA_ARGS__); break;     \
    case 33: hipLaunchKernelGGL((KERNEL<T, 16, true>), grid, dim3(64), 0, s, __VA_ARGS__); break;      \
    case 48: hipLaunchKernelGGL((KERNEL<T, 24, false>), grid, dim3(64), 0, s, __VA_ARGS__); break;     \
    case 49: hipLaunchKernelGGL((KERNEL<T, 24, true>), grid, dim3(64), 0, s, __VA_ARGS__); break;      \
    case 64: hipLaunchKernelGGL((KERNEL<T, 32, false>), grid, dim3(64), 0, s, __VA_ARGS__); break;     \
    case 65: hipLaunchKernelGGL((KERNEL<T, 32, true>), grid, dim3(64), 0, s, __VA_ARGS__); break;      \
    case 96: hipLaunchKernelGGL((KERNEL<T, 48, false>), grid, dim3(64), 0, s, __VA_ARGS__); break;     \
    case 97: hipLaunchKernelGGL((KERNEL<T, 48, true>), grid, dim3(64), 0, s, __VA_ARGS__); break;      \
    case 128: hipLaunchKernelGGL((KERNEL<T, 64, false>), grid, dim3(64), 0, s, __VA_ARGS__); break;    \
    default: hipLaunchKernelGGL((KERNEL<T, 64, true>), grid, dim3(64), 0, s, __VA_ARGS__); break;      \
  }

template <class T>
hipError_t launch_rollout_step(const ModelTable<T> *tab, const ModelTable<double> *tab64, int nj, int64_t B, T dt,
                               T *q, T *qd, const T *tau, int64_t tau_stride, T *save, int64_t save_stride,
                               hipStream_t s) {
  if (B <= 0) return hipSuccess;
  dim3 grid((unsigned)((B + 63) / 64));
  SAPAMD_ROLLOUT_DISPATCH(k_rollout_step, tab, tab64, B, dt, q, qd, tau, tau_stride, save, save_stride);
  return hipGetLastError();
}

template <class T>
hipError_t launch_rollout_step_vjp(const ModelTable<T> *tab, const ModelTable<double> *tab64, int nj, int64_t B,
                                   T dt, const T *state, int64_t state_stride, const T *tau, int64_t tau_stride,
                                   T *aq, T *aqd, T *gtau, hipStream_t s) {
  if (B <= 0) return hipSuccess;
  dim3 grid((unsigned)((B + 63) / 64));
  SAPAMD_ROLLOUT_DISPATCH(k_rollout_step_vjp, tab, tab64, B, dt, state, state_stride, tau, tau_stride, aq, aqd,
                          gtau);
  return hipGetLastError();
}

hipError_t launch_fd_ldlt(const ModelTable<double> *tab64, int nj, int64_t B, const float *q, const float *qd,
                          const float *tau, float *qdd, hipStream_t s) {
  if (B <= 0) return hipSuccess;
  dim3 grid((unsigned)((B + 63) / 64));
  switch (pick(nj)) {
    case 8: hipLaunchKernelGGL((k_fd_ldlt<8>), grid, dim3(64), 0, s, tab64, B, q, qd, tau, qdd); break;
    case 16: hipLaunchKernelGGL((k_fd_ldlt<16>), grid, dim3(64), 0, s, tab64, B, q, qd, tau, qdd); break;
    case 24: hipLaunchKernelGGL((k_fd_ldlt<24>), grid, dim3(64), 0, s, tab64, B, q, qd, tau, qdd); break;
    case 48: hipLaunchKernelGGL((k_fd_ldlt<48>), grid, dim3(64), 0, s, tab64, B, q, qd, tau, qdd); break;
    case 32: hipLaunchKernelGGL((k_fd_ldlt<32>), grid, dim3(64), 0, s, tab64, B, q, qd, tau, qdd); break;
    default: hipLaunchKernelGGL((k_fd_ldlt<64>), grid, dim3(64), 0, s, tab64, B, q, qd, tau, qdd); break;
  }
  return hipGetLastError();
}

template hipError_t launch_rollout_step<float>(const ModelTable<float> *, const ModelTable<double> *, int, int64_t,
                                               float, float *, float *, const float *, int64_t, float *, int64_t,
                                               hipStream_t);
template hipError_t launch_rollout_step_vjp<float>(const ModelTable<float> *, const ModelTable<double> *, int,
                                                   int64_t, float, const float *, int64_t, const float *, int64_t,
                                                   float *, float *, float *, hipStream_t);

}  // namespace sapamd
