// Reverse-mode kernels (vector-Jacobian products) of FK link poses and RNEA,
// thread-per-environment, any tree, FP32 / FP64.
//
// World-frame adjoint of RNEA. With the upstream gradient g on tau, define the
// virtual velocity w_j = sum_{i ancestor-or-self of j} g_i col_i, so that
// g . tau = sum_j w_j . f_j. Differentiating this virtual power with every
// body-attached vector moving with the joints above it gives, per joint k
// (all sums over the subtree of k, parent quantities marked lam):
//   Hw_k = sum I_j w_j
//   QS_k = sum [ -v_j x* (I_j w_j) + w_j x* (I_j v_j) + I_j (v_j x w_j) ]
//   d/dqdd_k = col_k . Hw_k
//   d/dqd_k  = -col_k . (QS_k + 2 v_lam x* Hw_k)
//   d/dq_k   = col_k . ( -w_lam x* F_k - a_lam x* Hw_k + v_lam x* (QS_k + v_lam x* Hw_k) )
// where F_k is the RNEA subtree force and w_lam = w_k - g_k col_k,
// v_lam = v_k - qd_k col_k, a_lam = a_k - qdd_k col_k - qd_k (v_k x col_k) are
// recovered from the joint's own values.
//
// FK link poses: an upstream gradient (gp, gq) on a link's position and
// quaternion (w, u) acts as the world wrench [gp; p x gp + tq] with
// tq = 1/2 (w gq_u - gq_w u + u x gq_u) (from qdot = 1/2 (0, omega) (x) q);
// d/dq_k = col_k . (sum of the wrenches of the links in the subtree of k).
// Both are checked against exact dual-number derivatives of the FP64 oracle.
#include <hip/hip_runtime.h>

#include "launch.h"
#include "spatial.h"

namespace sapamd {

namespace {

template <class T> __device__ __forceinline__ S6<T> link_wrench(V3<T> p, const T *quat, const T *g) {
  V3<T> gp = {g[0], g[1], g[2]};
  const T w = quat[0];
  V3<T> u = {quat[1], quat[2], quat[3]}, gu = {g[4], g[5], g[6]};
  V3<T> tq = T(0.5) * (w * gu - g[3] * u + cross(u, gu));
  return {gp, cross(p, gp) + tq};
}

template <class T, int MJ>
__global__ __launch_bounds__(64) void k_kin_vjp(const ModelTable<T> *__restrict__ M, int64_t B,
                                                const T *__restrict__ q, const T *__restrict__ qd,
                                                const T *__restrict__ qdd, VjpArgs<T> A) {
  int64_t b = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (b >= B) return;
  const int nj = M->nj, nv = M->nv, L = M->nlinks;
  const JointRec<T> *JR = joints_of(M);
  const LinkRec<T> *LR = links_of(M);
  const bool dyn = A.gtau || A.tau_star;
  const bool lossmode = A.tau_star || A.p_target;
  T oR[MJ][9];
  V3<T> op[MJ];
  S6<T> col[MJ], vel[MJ], acc[MJ], F[MJ], W[MJ], Hw[MJ], QS[MJ], FP[MJ];
  T gt[MJ];
  const S6<T> a0 = {{-M->gravity[0], -M->gravity[1], -M->gravity[2]}, {T(0), T(0), T(0)}};
  const T *qb = q + b * nv;
  for (int j = 0; j < nj; ++j) {
    const JointRec<T> &J = JR[j];
    const int u = J.vuser, p = J.parent;
    T qj = qb[u], s, c;
    sincos_t(qj, &s, &c);
    joint_world(J, qj, s, c, p >= 0 ? oR[p] : static_cast<const T *>(nullptr), p >= 0 ? op[p] : V3<T>{0, 0, 0},
                p < 0, oR[j], op[j]);
    col[j] = jcol(J, oR[j], op[j]);
    FP[j] = zero6<T>();
    if (dyn) {
      const T qdj = qd[b * nv + u], qddj = qdd[b * nv + u];
      vel[j] = (p >= 0 ? vel[p] : zero6<T>()) + qdj * col[j];
      acc[j] = (p >= 0 ? acc[p] : a0) + qddj * col[j] + qdj * crossm(vel[j], col[j]);
      const WInertia<T> Y = world_inertia(J, oR[j], op[j]);
      F[j] = inertia_apply(Y, acc[j]) + crossf(vel[j], inertia_apply(Y, vel[j]));
    }
  }
  T loss = T(0);
  if (dyn) {
    for (int j = nj - 1; j >= 0; --j) {
      const int p = JR[j].parent;
      if (p >= 0) F[p] = F[p] + F[j];
    }
    for (int j = 0; j < nj; ++j) {
      const int u = JR[j].vuser;
      const T tj = dot6(col[j], F[j]);
      if (A.tau) A.tau[b * nv + u] = tj;
      if (A.tau_star) {
        const T d = tj - A.tau_star[b * nv + u];
        loss += d * d;
        gt[j] = T(2) * d;
      } else {
        gt[j] = A.gtau[b * nv + u];
      }
    }
  }
  // link-pose wrenches
  if (A.gposes || A.p_target) {
    for (int l = 0; l < L; ++l) {
      const LinkRec<T> &Lk = LR[l];
      const int jl = Lk.joint;
      const bool want = A.gposes ? true : l == A.ee_link;
      if (jl < 0 || !want) continue;
      T R[9];
      mm(oR[jl], Lk.R, R);
      const V3<T> pw = mv(oR[jl], load3(Lk.p)) + op[jl];
      if (A.gposes) {
        T quat[4];
        quat_of(R, quat);
        FP[jl] = FP[jl] + link_wrench(pw, quat, A.gposes + (b * L + l) * 7);
      } else {
        const V3<T> d = pw - V3<T>{A.p_target[b * 3], A.p_target[b * 3 + 1], A.p_target[b * 3 + 2]};
        loss += dot(d, d);
        const V3<T> gp = T(2) * d;
        FP[jl] = FP[jl] + S6<T>{gp, cross(pw, gp)};
      }
    }
  }
  if (dyn) {
    for (int j = 0; j < nj; ++j) {
      const JointRec<T> &J = JR[j];
      const int p = J.parent;
      W[j] = (p >= 0 ? W[p] : zero6<T>()) + gt[j] * col[j];
      const WInertia<T> Y = world_inertia(J, oR[j], op[j]);
      const S6<T> Iw = inertia_apply(Y, W[j]), Iv = inertia_apply(Y, vel[j]);
      Hw[j] = Iw;
      QS[j] = crossf(W[j], Iv) - crossf(vel[j], Iw) + inertia_apply(Y, crossm(vel[j], W[j]));
    }
  }
  for (int j = nj - 1; j >= 0; --j) {
    const int p = JR[j].parent;
    if (p < 0) continue;
    FP[p] = FP[p] + FP[j];
    if (dyn) {
      Hw[p] = Hw[p] + Hw[j];
      QS[p] = QS[p] + QS[j];
    }
  }
  for (int k = 0; k < nj; ++k) {
    const int u = JR[k].vuser;
    const S6<T> c = col[k];
    T dq = dot6(c, FP[k]);
    if (dyn) {
      const T qdk = qd[b * nv + u], qddk = qdd[b * nv + u];
      const S6<T> wl = W[k] - gt[k] * c, vl = vel[k] - qdk * c;
      const S6<T> al = acc[k] - qddk * c - qdk * crossm(vel[k], c);
      dq += dot6(c, crossf(vl, QS[k] + crossf(vl, Hw[k])) - crossf(wl, F[k]) - crossf(al, Hw[k]));
      if (A.dqd) A.dqd[b * nv + u] = -dot6(c, QS[k] + T(2) * crossf(vl, Hw[k]));
      if (A.dqdd) A.dqdd[b * nv + u] = dot6(c, Hw[k]);
    }
    if (A.dq) A.dq[b * nv + u] = dq;
  }
  if (lossmode && A.loss) A.loss[b] = loss;
}

inline int pick_mj(int nj) {
  return nj <= 8 ? 8 : nj <= 16 ? 16 : nj <= 24 ? 24 : nj <= 32 ? 32 : nj <= 48 ? 48 : 64;
}

}  // namespace

template <class T>
hipError_t launch_kin_vjp(const ModelTable<T> *tab, int nj, int64_t B, const T *q, const T *qd, const T *qdd,
                          VjpArgs<T> a, hipStream_t s) {
  if (B <= 0) return hipSuccess;
  dim3 grid((unsigned)((B + 63) / 64));
  switch (pick_mj(nj)) {
    case 8: hipLaunchKernelGGL((k_kin_vjp<T, 8>), grid, dim3(64), 0, s, tab, B, q, qd, qdd, a); break;
    case 16: hipLaunchKernelGGL((k_kin_vjp<T, 16>), grid, dim3(64), 0, s, tab, B, q, qd, qdd, a); break;
    case 24: hipLaunchKernelGGL((k_kin_vjp<T, 24>), grid, dim3(64), 0, s, tab, B, q, qd, qdd, a); break;
    case 32: hipLaunchKernelGGL((k_kin_vjp<T, 32>), grid, dim3(64), 0, s, tab, B, q, qd, qdd, a); break;
    case 48: hipLaunchKernelGGL((k_kin_vjp<T, 48>), grid, dim3(64), 0, s, tab, B, q, qd, qdd, a); break;
    default: hipLaunchKernelGGL((k_kin_vjp<T, 64>), grid, dim3(64), 0, s, tab, B, q, qd, qdd, a); break;
  }
  return hipGetLastError();
}

template hipError_t launch_kin_vjp<float>(const ModelTable<float> *, int, int64_t, const float *, const float *,
                                          const float *, VjpArgs<float>, hipStream_t);
template hipError_t launch_kin_vjp<double>(const ModelTable<double> *, int, int64_t, const double *, const double *,
                                           const double *, VjpArgs<double>, hipStream_t);

}  // namespace sapamd
