// Generic thread-per-environment kernels for any tree up to 64 joints, in FP32
// (batched path, heterogeneous fallback) and FP64 (the drop-in per-call path).
//
// Each lane owns one environment and sweeps the joints in topological order with
// the world-frame recursions of spatial.h:
//   FK        oMi[j] = oMi[parent] * place_j * X_J(q_j)            (pinocchio_model.cpp:146-148)
//   Jacobian  col_j = [p_j x a_j; a_j] or [a_j; 0]                 (:163-186)
//   RNEA      v_j = v_par + col_j qd_j, a_j = a_par + col_j qdd_j + (v_j x col_j) qd_j,
//             f_j = I_j a_j + v_j x* I_j v_j, F_par += F_j, tau_j = col_j . F_j   (:214-219)
//   CRBA      Yc_par += Yc_j, M_ij = col_i . (Yc_j col_j), i ancestor-or-self of j   (:201-206)
//   ABA       Featherstone's three passes in the world frame (no frame changes)       (:221-226)
//   Coriolis  C = sum_i J_i^T (I_i Jdot_i + B(I_i, v_i) J_i) by subtree sums (Christoffel-consistent) (:208-212)
//   CLIK      damped least squares on log6 with Jlog6, per-problem early exit        (:228-300)
// Per-joint state lives in private arrays sized by the next power of two >= nj.
#include <hip/hip_runtime.h>

#include "launch.h"
#include "se3.h"
#include "spatial.h"

namespace sapamd {

namespace {

template <class T> __device__ __forceinline__ S6<T> gravity_acc(const ModelTable<T> *M) {
  return {{-M->gravity[0], -M->gravity[1], -M->gravity[2]}, {T(0), T(0), T(0)}};
}

// Per-joint arrays of the thread-per-environment kernels. LDS = false: private
// arrays (scratch), one environment per lane. LDS = true: the arrays live in
// the block's LDS, for one-environment launches (the per-call path, one
// thread per block): a private segment is reserved per hardware queue for good
// from an agent-wide budget (hsa_ext_amd.h, SCRATCH_LIMIT_MAX), and FP64 arrays
// of 17-64 joints would reserve tens of GB of it.
#define SAPAMD_JARR(TYPE, NAME)            \
  __shared__ TYPE NAME##_s[LDS ? MJ : 1]; \
  TYPE NAME##_l[LDS ? 1 : MJ];            \
  TYPE *const NAME = LDS ? NAME##_s : NAME##_l;


// One-environment launches (LDS arrays shared by the block): world pose of
// every joint. Each thread takes joints for their own transforms (sine,
// cosine, placement), then thread 0 composes them root to leaves. Every
// thread of the block calls it.
template <class T>
__device__ __forceinline__ void shared_joint_world(const ModelTable<T> *M, const T *qb, int tid, int nth, T (*oR)[9],
                                                   V3<T> *op, int *par) {
  const int nj = M->nj;
  const JointRec<T> *JR = joints_of(M);
  for (int j = tid; j < nj; j += nth) {
    const JointRec<T> &J = JR[j];
    T qj = qb[J.vuser], s, c;
    sincos_t(qj, &s, &c);
    joint_world(J, qj, s, c, static_cast<const T *>(nullptr), V3<T>{0, 0, 0}, true, oR[j], op[j]);
    par[j] = J.parent;
  }
  __syncthreads();
  if (tid == 0)
    for (int j = 0; j < nj; ++j) {
      const int p = par[j];
      if (p < 0) continue;
      T R[9];
      mm(oR[p], oR[j], R);
      op[j] = mv(oR[p], op[j]) + op[p];
#pragma unroll
      for (int k = 0; k < 9; ++k) oR[j][k] = R[k];
    }
  __syncthreads();
}

// LDS = true (one environment, the per-call path): the block's threads share
// the per-joint arrays. Per-joint work (the joint's own transform with its
// sine and cosine, its Jacobian column, its spatial force) runs one joint per
// thread; thread 0 runs only the root-to-leaf compositions and the leaf-to-root
// force sums between them; then every thread takes links, and (link, dof)
// entries of the per-link Jacobians.
template <class T, int MJ, bool LDS>
__global__ __launch_bounds__(64) void k_kinematics(const ModelTable<T> *__restrict__ M, int64_t B, int flags,
                                                   const T *__restrict__ q, const T *__restrict__ qd,
                                                   const T *__restrict__ qdd, KinOut<T> out) {
  const int64_t b = LDS ? (int64_t)blockIdx.x : blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int tid = LDS ? (int)threadIdx.x : 0, nth = LDS ? (int)blockDim.x : 1;
  if (b >= B) return;
  const int nj = M->nj, nv = M->nv, L = M->nlinks;
  const JointRec<T> *JR = joints_of(M);
  const LinkRec<T> *LR = links_of(M);
  using R9 = T[9];
  SAPAMD_JARR(R9, oR)
  SAPAMD_JARR(V3<T>, op)
  SAPAMD_JARR(S6<T>, col)
  SAPAMD_JARR(S6<T>, vel)
  SAPAMD_JARR(S6<T>, acc)
  SAPAMD_JARR(S6<T>, F)
  const bool dyn = flags & KF_RNEA;
  const S6<T> a0 = (flags & KF_NOGRAV) ? zero6<T>() : gravity_acc(M);
  const T *qb = q + b * (out.ld_q ? out.ld_q : nv);
  const T *qdb = qd + b * (out.ld_qd ? out.ld_qd : nv), *qddb = qdd + b * (out.ld_qdd ? out.ld_qdd : nv);
  SAPAMD_JARR(int, joint_of_dof)
  if constexpr (LDS) {
    SAPAMD_JARR(int, par)
    SAPAMD_JARR(T, qdv)
    SAPAMD_JARR(T, qddv)
    // every input read in one pass (q with the joints' own transforms)
    for (int j = tid; j < nj; j += nth) {
      qdv[j] = dyn && !(flags & KF_QD0) ? qdb[JR[j].vuser] : T(0);
      qddv[j] = dyn && !(flags & KF_QDD0) ? qddb[JR[j].vuser] : T(0);
      joint_of_dof[JR[j].vuser] = j;
    }
    shared_joint_world(M, qb, tid, nth, oR, op, par);
    for (int j = tid; j < nj; j += nth) {
      col[j] = jcol(JR[j], oR[j], op[j]);
      if (dyn) acc[j] = qddv[j] * col[j];
    }
    if (dyn) {
      __syncthreads();
      if (tid == 0)  // velocities and accelerations, root to leaves
        for (int j = 0; j < nj; ++j) {
          const int p = par[j];
          const S6<T> vp = p >= 0 ? vel[p] : zero6<T>(), ap = p >= 0 ? acc[p] : a0;
          vel[j] = vp + qdv[j] * col[j];
          acc[j] = ap + acc[j] + qdv[j] * crossm(vel[j], col[j]);
        }
      __syncthreads();
      for (int j = tid; j < nj; j += nth) {
        const WInertia<T> Y = world_inertia(JR[j], oR[j], op[j]);
        F[j] = inertia_apply(Y, acc[j]) + crossf(vel[j], inertia_apply(Y, vel[j]));
      }
      __syncthreads();
      if (tid == 0)  // forces, leaves to root
        for (int j = nj - 1; j >= 0; --j) {
          out.tau[b * (out.ld_tau ? out.ld_tau : nv) + JR[j].vuser] = dot6(col[j], F[j]);
          const int p = par[j];
          if (p >= 0) F[p] = F[p] + F[j];
        }
    }
  } else {
  for (int j = 0; j < nj; ++j) {
    const JointRec<T> &J = JR[j];
    const int u = J.vuser, p = J.parent;
    T qj = qb[u], s, c;
    sincos_t(qj, &s, &c);
    joint_world(J, qj, s, c, p >= 0 ? oR[p] : nullptr, p >= 0 ? op[p] : V3<T>{0, 0, 0}, p < 0, oR[j], op[j]);
    col[j] = jcol(J, oR[j], op[j]);
    if (dyn) {
      T qdj = (flags & KF_QD0) ? T(0) : qdb[u], qddj = (flags & KF_QDD0) ? T(0) : qddb[u];
      S6<T> vp = p >= 0 ? vel[p] : zero6<T>();
      S6<T> ap = p >= 0 ? acc[p] : a0;
      vel[j] = vp + qdj * col[j];
      acc[j] = ap + qddj * col[j] + qdj * crossm(vel[j], col[j]);
      WInertia<T> Y = world_inertia(J, oR[j], op[j]);
      F[j] = inertia_apply(Y, acc[j]) + crossf(vel[j], inertia_apply(Y, vel[j]));
    }
  }
  if (dyn) {
    for (int j = nj - 1; j >= 0; --j) {
      out.tau[b * (out.ld_tau ? out.ld_tau : nv) + JR[j].vuser] = dot6(col[j], F[j]);
      int p = JR[j].parent;
      if (p >= 0) F[p] = F[p] + F[j];
    }
  }
  }  // LDS
  if (LDS) __syncthreads();
  if (flags & KF_JFULL) {
    T *Jb = out.jfull + b * 6 * nv;
    for (int j = tid; j < nj; j += nth) {
      int u = JR[j].vuser;
      Jb[0 * nv + u] = col[j].l.x; Jb[1 * nv + u] = col[j].l.y; Jb[2 * nv + u] = col[j].l.z;
      Jb[3 * nv + u] = col[j].a.x; Jb[4 * nv + u] = col[j].a.y; Jb[5 * nv + u] = col[j].a.z;
    }
  }
  if (LDS && (flags & KF_JLINKS)) {
    // entry (l, u): the column of the joint on user dof u (world frame, and in
    // the link frame about its origin) when that joint supports link l, else 0
    for (int it = tid; it < L * nv; it += nth) {
      const int l = it / nv, u = it % nv;
      const LinkRec<T> &Lk = LR[l];
      const int jl = Lk.joint;
      const int j = joint_of_dof[u];
      const bool anc = jl >= 0 && j <= jl && jl < j + JR[j].subtree;
      S6<T> cw = zero6<T>(), cl = zero6<T>();
      if (anc) {
        T R[9];
        mm(oR[jl], Lk.R, R);
        const V3<T> pw = mv(oR[jl], load3(Lk.p)) + op[jl];
        cw = col[j];
        cl = to_local(R, pw, cw);
      }
      T *Jw = out.jw_links + (b * L + l) * 6 * nv + u;
      T *Jl = out.jl_links + (b * L + l) * 6 * nv + u;
      Jw[0 * nv] = cw.l.x; Jw[1 * nv] = cw.l.y; Jw[2 * nv] = cw.l.z;
      Jw[3 * nv] = cw.a.x; Jw[4 * nv] = cw.a.y; Jw[5 * nv] = cw.a.z;
      Jl[0 * nv] = cl.l.x; Jl[1 * nv] = cl.l.y; Jl[2 * nv] = cl.l.z;
      Jl[3 * nv] = cl.a.x; Jl[4 * nv] = cl.a.y; Jl[5 * nv] = cl.a.z;
    }
    if (!(flags & KF_POSES)) return;
  }
  if (flags & (KF_POSES | KF_JLINKS)) {
    for (int l = tid; l < L; l += nth) {
      const LinkRec<T> &Lk = LR[l];
      const int jl = Lk.joint;
      T R[9];
      V3<T> pw;
      if (jl < 0) {
        for (int k = 0; k < 9; ++k) R[k] = Lk.R[k];
        pw = load3(Lk.p);
      } else {
        mm(oR[jl], Lk.R, R);
        pw = mv(oR[jl], load3(Lk.p)) + op[jl];
      }
      if (flags & KF_POSES) {
        T *P = out.poses + b * (out.pose_env ? out.pose_env : (int64_t)L * 7) + l * (out.pose_link ? out.pose_link : 7);
        P[0] = pw.x; P[1] = pw.y; P[2] = pw.z;
        quat_of(R, P + 3);
      }
      if (!LDS && (flags & KF_JLINKS)) {
        T *Jw = out.jw_links + (b * L + l) * 6 * nv;
        T *Jl = out.jl_links + (b * L + l) * 6 * nv;
        for (int k = 0; k < 6 * nv; ++k) { Jw[k] = T(0); Jl[k] = T(0); }
        for (int k = jl; k >= 0; k = JR[k].parent) {
          int u = JR[k].vuser;
          S6<T> cw = col[k], cl = to_local(R, pw, cw);
          Jw[0 * nv + u] = cw.l.x; Jw[1 * nv + u] = cw.l.y; Jw[2 * nv + u] = cw.l.z;
          Jw[3 * nv + u] = cw.a.x; Jw[4 * nv + u] = cw.a.y; Jw[5 * nv + u] = cw.a.z;
          Jl[0 * nv + u] = cl.l.x; Jl[1 * nv + u] = cl.l.y; Jl[2 * nv + u] = cl.l.z;
          Jl[3 * nv + u] = cl.a.x; Jl[4 * nv + u] = cl.a.y; Jl[5 * nv + u] = cl.a.z;
        }
      }
    }
  }
}

template <class T, int MJ, bool LDS>
__global__ __launch_bounds__(64) void k_link_local_jacobian(const ModelTable<T> *__restrict__ M, int64_t B, int link,
                                                            const T *__restrict__ q, T *__restrict__ Jout) {
  int64_t b = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (b >= B) return;
  const int nj = M->nj, nv = M->nv;
  const JointRec<T> *JR = joints_of(M);
  const LinkRec<T> &Lk = links_of(M)[link];
  using R9 = T[9];
  SAPAMD_JARR(R9, oR)
  SAPAMD_JARR(V3<T>, op)
  const T *qb = q + b * nv;
  for (int j = 0; j < nj; ++j) {
    const JointRec<T> &J = JR[j];
    int p = J.parent;
    T qj = qb[J.vuser], s, c;
    sincos_t(qj, &s, &c);
    joint_world(J, qj, s, c, p >= 0 ? oR[p] : nullptr, p >= 0 ? op[p] : V3<T>{0, 0, 0}, p < 0, oR[j], op[j]);
  }
  T *Jb = Jout + b * 6 * nv;
  for (int k = 0; k < 6 * nv; ++k) Jb[k] = T(0);
  const int jl = Lk.joint;
  if (jl < 0) return;
  T R[9];
  mm(oR[jl], Lk.R, R);
  V3<T> pw = mv(oR[jl], load3(Lk.p)) + op[jl];
  for (int k = jl; k >= 0; k = JR[k].parent) {
    int u = JR[k].vuser;
    S6<T> cl = to_local(R, pw, jcol(JR[k], oR[k], op[k]));
    Jb[0 * nv + u] = cl.l.x; Jb[1 * nv + u] = cl.l.y; Jb[2 * nv + u] = cl.l.z;
    Jb[3 * nv + u] = cl.a.x; Jb[4 * nv + u] = cl.a.y; Jb[5 * nv + u] = cl.a.z;
  }
}

template <class T, int MJ, bool LDS>
__global__ __launch_bounds__(64) void k_crba(const ModelTable<T> *__restrict__ M, int64_t B, const T *__restrict__ q,
                                             T *__restrict__ Mout) {
  const int64_t b = LDS ? (int64_t)blockIdx.x : blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (b >= B) return;
  const int nj = M->nj, nv = M->nv;
  const JointRec<T> *JR = joints_of(M);
  using R9 = T[9];
  SAPAMD_JARR(R9, oR)
  SAPAMD_JARR(V3<T>, op)
  SAPAMD_JARR(S6<T>, col)
  SAPAMD_JARR(AInertia<T>, Yc)
  const T *qb = q + b * nv;
  if constexpr (LDS) {
    // one environment, the block's threads: per-joint work one joint per
    // thread, the composite sums on thread 0, then one thread per entry of M
    const int tid = threadIdx.x, nth = blockDim.x;
    SAPAMD_JARR(int, par)
    SAPAMD_JARR(int, joint_of_dof)
    SAPAMD_JARR(S6<T>, F)
    shared_joint_world(M, qb, tid, nth, oR, op, par);
    for (int j = tid; j < nj; j += nth) {
      col[j] = jcol(JR[j], oR[j], op[j]);
      Yc[j] = additive(world_inertia(JR[j], oR[j], op[j]));
      joint_of_dof[JR[j].vuser] = j;
    }
    __syncthreads();
    if (tid == 0)
      for (int j = nj - 1; j > 0; --j) {
        const int p = par[j];
        if (p >= 0) add_to(Yc[p], Yc[j]);
      }
    __syncthreads();
    for (int j = tid; j < nj; j += nth) F[j] = apply(Yc[j], col[j]);
    __syncthreads();
    T *Mb = Mout + b * nv * nv;
    for (int e = tid; e < nv * nv; e += nth) {
      const int i = joint_of_dof[e / nv], j = joint_of_dof[e % nv];
      // M_ij = col_a . F_d with a the ancestor-or-self, d the descendant
      const int a = i < j ? i : j, d = i < j ? j : i;
      Mb[e] = (a <= d && d < a + JR[a].subtree) ? dot6(col[a], F[d]) : T(0);
    }
    return;
  }
  for (int j = 0; j < nj; ++j) {
    const JointRec<T> &J = JR[j];
    int p = J.parent;
    T qj = qb[J.vuser], s, c;
    sincos_t(qj, &s, &c);
    joint_world(J, qj, s, c, p >= 0 ? oR[p] : nullptr, p >= 0 ? op[p] : V3<T>{0, 0, 0}, p < 0, oR[j], op[j]);
    col[j] = jcol(J, oR[j], op[j]);
    Yc[j] = additive(world_inertia(J, oR[j], op[j]));
  }
  for (int j = nj - 1; j > 0; --j) {
    int p = JR[j].parent;
    if (p >= 0) add_to(Yc[p], Yc[j]);
  }
  T *Mb = Mout + b * nv * nv;
  for (int k = 0; k < nv * nv; ++k) Mb[k] = T(0);
  for (int j = 0; j < nj; ++j) {
    S6<T> Fj = apply(Yc[j], col[j]);
    int uj = JR[j].vuser;
    for (int i = j; i >= 0; i = JR[i].parent) {
      int ui = JR[i].vuser;
      T v = dot6(col[i], Fj);
      Mb[ui * nv + uj] = v;
      Mb[uj * nv + ui] = v;
    }
  }
}

// ---- dense 6x6 helpers (row-major) for ABA and the Coriolis matrix
template <class T> struct Mat6 {
  T a[36];
};
template <class T> __device__ __forceinline__ void skew_into(T *m, int r0, int c0, V3<T> v, T sgn) {
  m[(r0 + 0) * 6 + c0 + 0] = T(0);     m[(r0 + 0) * 6 + c0 + 1] = -sgn * v.z; m[(r0 + 0) * 6 + c0 + 2] = sgn * v.y;
  m[(r0 + 1) * 6 + c0 + 0] = sgn * v.z; m[(r0 + 1) * 6 + c0 + 1] = T(0);     m[(r0 + 1) * 6 + c0 + 2] = -sgn * v.x;
  m[(r0 + 2) * 6 + c0 + 0] = -sgn * v.y; m[(r0 + 2) * 6 + c0 + 1] = sgn * v.x; m[(r0 + 2) * 6 + c0 + 2] = T(0);
}
// Y as a 6x6 matrix: [[m I, -[h]x], [[h]x, I_O]].
template <class T> __device__ __forceinline__ Mat6<T> mat_of(const AInertia<T> &Y) {
  Mat6<T> m;
  for (int k = 0; k < 36; ++k) m.a[k] = T(0);
  m.a[0] = m.a[7] = m.a[14] = Y.m;
  skew_into(m.a, 0, 3, Y.h, T(-1));
  skew_into(m.a, 3, 0, Y.h, T(1));
  m.a[21] = Y.I[0]; m.a[22] = Y.I[1]; m.a[23] = Y.I[2];
  m.a[27] = Y.I[1]; m.a[28] = Y.I[3]; m.a[29] = Y.I[4];
  m.a[33] = Y.I[2]; m.a[34] = Y.I[4]; m.a[35] = Y.I[5];
  return m;
}
template <class T> __device__ __forceinline__ S6<T> mul6(const Mat6<T> &A, S6<T> x) {
  T v[6] = {x.l.x, x.l.y, x.l.z, x.a.x, x.a.y, x.a.z}, r[6];
  for (int i = 0; i < 6; ++i) {
    T acc = T(0);
    for (int k = 0; k < 6; ++k) acc += A.a[i * 6 + k] * v[k];
    r[i] = acc;
  }
  return {{r[0], r[1], r[2]}, {r[3], r[4], r[5]}};
}
template <class T> __device__ __forceinline__ S6<T> mul6t(const Mat6<T> &A, S6<T> x) {
  T v[6] = {x.l.x, x.l.y, x.l.z, x.a.x, x.a.y, x.a.z}, r[6];
  for (int i = 0; i < 6; ++i) {
    T acc = T(0);
    for (int k = 0; k < 6; ++k) acc += A.a[k * 6 + i] * v[k];
    r[i] = acc;
  }
  return {{r[0], r[1], r[2]}, {r[3], r[4], r[5]}};
}

template <class T, int MJ, bool LDS>
__global__ __launch_bounds__(64) void k_aba(const ModelTable<T> *__restrict__ M, int64_t B, const T *__restrict__ q,
                                            const T *__restrict__ qd, const T *__restrict__ tau,
                                            T *__restrict__ qdd) {
  const int64_t b = LDS ? (int64_t)blockIdx.x : blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (b >= B) return;
  const int nj = M->nj, nv = M->nv;
  const JointRec<T> *JR = joints_of(M);
  using R9 = T[9];
  SAPAMD_JARR(R9, oR)
  SAPAMD_JARR(V3<T>, op)
  SAPAMD_JARR(S6<T>, col)
  SAPAMD_JARR(S6<T>, vel)
  SAPAMD_JARR(S6<T>, cb)
  SAPAMD_JARR(S6<T>, pA)
  SAPAMD_JARR(S6<T>, U)
  SAPAMD_JARR(Mat6<T>, IA)
  SAPAMD_JARR(T, D)
  SAPAMD_JARR(T, uu)
  SAPAMD_JARR(T, tauv)  // the per-call path's torques, read once up front
  const T *qb = q + b * nv;
  if constexpr (LDS) {
    // one environment, the block's threads: per-joint terms one joint per
    // thread; the velocity prefix and the two recursions below on thread 0
    const int tid = threadIdx.x, nth = blockDim.x;
    SAPAMD_JARR(int, par)
    SAPAMD_JARR(T, qdv)
    // every input read in one pass (q with the joints' own transforms)
    for (int j = tid; j < nj; j += nth) {
      qdv[j] = qd[b * nv + JR[j].vuser];
      tauv[j] = tau[b * nv + JR[j].vuser];
    }
    shared_joint_world(M, qb, tid, nth, oR, op, par);
    for (int j = tid; j < nj; j += nth) col[j] = jcol(JR[j], oR[j], op[j]);
    __syncthreads();
    if (tid == 0)
      for (int j = 0; j < nj; ++j) vel[j] = (par[j] >= 0 ? vel[par[j]] : zero6<T>()) + qdv[j] * col[j];
    __syncthreads();
    for (int j = tid; j < nj; j += nth) {
      cb[j] = qdv[j] * crossm(vel[j], col[j]);
      const WInertia<T> Y = world_inertia(JR[j], oR[j], op[j]);
      IA[j] = mat_of(additive(Y));
      pA[j] = crossf(vel[j], inertia_apply(Y, vel[j]));
    }
    __syncthreads();
    if (tid != 0) return;
  } else {
  for (int j = 0; j < nj; ++j) {
    const JointRec<T> &J = JR[j];
    int p = J.parent, u = J.vuser;
    T qj = qb[u], s, c;
    sincos_t(qj, &s, &c);
    joint_world(J, qj, s, c, p >= 0 ? oR[p] : nullptr, p >= 0 ? op[p] : V3<T>{0, 0, 0}, p < 0, oR[j], op[j]);
    col[j] = jcol(J, oR[j], op[j]);
    T qdj = qd[b * nv + u];
    vel[j] = (p >= 0 ? vel[p] : zero6<T>()) + qdj * col[j];
    cb[j] = qdj * crossm(vel[j], col[j]);
    WInertia<T> Y = world_inertia(J, oR[j], op[j]);
    IA[j] = mat_of(additive(Y));
    pA[j] = crossf(vel[j], inertia_apply(Y, vel[j]));
  }
  }  // LDS
  for (int j = nj - 1; j >= 0; --j) {
    U[j] = mul6(IA[j], col[j]);
    D[j] = dot6(col[j], U[j]);
    uu[j] = (LDS ? tauv[j] : tau[b * nv + JR[j].vuser]) - dot6(col[j], pA[j]);
    int p = JR[j].parent;
    if (p >= 0) {
      T Uv[6];
      vec_of(U[j], Uv);
      T invD = T(1) / D[j];
      Mat6<T> Ia;
      for (int r = 0; r < 6; ++r)
        for (int k = 0; k < 6; ++k) Ia.a[r * 6 + k] = IA[j].a[r * 6 + k] - Uv[r] * Uv[k] * invD;
      S6<T> pa = pA[j] + mul6(Ia, cb[j]) + (uu[j] * invD) * U[j];
      for (int k = 0; k < 36; ++k) IA[p].a[k] += Ia.a[k];
      pA[p] = pA[p] + pa;
    }
  }
  const S6<T> a0 = gravity_acc(M);
  SAPAMD_JARR(S6<T>, acc)
  for (int j = 0; j < nj; ++j) {
    int p = JR[j].parent;
    acc[j] = (p >= 0 ? acc[p] : a0) + cb[j];
    T x = (uu[j] - dot6(U[j], acc[j])) / D[j];
    qdd[b * nv + JR[j].vuser] = x;
    acc[j] = acc[j] + x * col[j];
  }
}

// B(I, v) = 1/2 [ (v x*) I + (I v) xbar* - I (v x) ]  (Echeandia & Wensing, 2021)
template <class T> __device__ Mat6<T> coriolis_body(const Mat6<T> &I, S6<T> v) {
  Mat6<T> crf, crm, bf, out;
  for (int k = 0; k < 36; ++k) { crf.a[k] = T(0); crm.a[k] = T(0); bf.a[k] = T(0); }
  // crm(v) = [[w]x, [v]x; 0, [w]x]; crf(v) = [[w]x, 0; [v]x, [w]x]
  skew_into(crm.a, 0, 0, v.a, T(1)); skew_into(crm.a, 0, 3, v.l, T(1)); skew_into(crm.a, 3, 3, v.a, T(1));
  skew_into(crf.a, 0, 0, v.a, T(1)); skew_into(crf.a, 3, 0, v.l, T(1)); skew_into(crf.a, 3, 3, v.a, T(1));
  // (f xbar*) m = m x* f, f = I v: [[0, -[fl]x], [-[fl]x, -[fa]x]]
  S6<T> f = mul6(I, v);
  skew_into(bf.a, 0, 3, f.l, T(-1)); skew_into(bf.a, 3, 0, f.l, T(-1)); skew_into(bf.a, 3, 3, f.a, T(-1));
  for (int r = 0; r < 6; ++r)
    for (int c = 0; c < 6; ++c) {
      T s1 = T(0), s2 = T(0);
      for (int k = 0; k < 6; ++k) {
        s1 += crf.a[r * 6 + k] * I.a[k * 6 + c];
        s2 += I.a[r * 6 + k] * crm.a[k * 6 + c];
      }
      out.a[r * 6 + c] = T(0.5) * (s1 + bf.a[r * 6 + c] - s2);
    }
  return out;
}

template <class T, int MJ, bool LDS>
__global__ __launch_bounds__(64) void k_coriolis(const ModelTable<T> *__restrict__ M, int64_t B,
                                                 const T *__restrict__ q, const T *__restrict__ qd,
                                                 T *__restrict__ Cout) {
  const int64_t b = LDS ? (int64_t)blockIdx.x : blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (b >= B) return;
  const int nj = M->nj, nv = M->nv;
  const JointRec<T> *JR = joints_of(M);
  using R9 = T[9];
  SAPAMD_JARR(R9, oR)
  SAPAMD_JARR(V3<T>, op)
  SAPAMD_JARR(S6<T>, col)
  SAPAMD_JARR(S6<T>, dcol)
  SAPAMD_JARR(S6<T>, vel)
  SAPAMD_JARR(Mat6<T>, Ic)
  SAPAMD_JARR(Mat6<T>, Bc)
  const T *qb = q + b * nv;
  if constexpr (LDS) {
    // one environment, the block's threads: per-joint terms one joint per
    // thread, the subtree sums one matrix entry per thread, then one thread
    // per entry of C
    const int tid = threadIdx.x, nth = blockDim.x;
    SAPAMD_JARR(int, par)
    SAPAMD_JARR(int, joint_of_dof)
    SAPAMD_JARR(T, qdv)
    SAPAMD_JARR(S6<T>, F1)
    SAPAMD_JARR(S6<T>, G)
    SAPAMD_JARR(S6<T>, H)
    for (int j = tid; j < nj; j += nth) {  // q' read with q
      qdv[j] = qd[b * nv + JR[j].vuser];
      joint_of_dof[JR[j].vuser] = j;
    }
    shared_joint_world(M, qb, tid, nth, oR, op, par);
    for (int j = tid; j < nj; j += nth) col[j] = jcol(JR[j], oR[j], op[j]);
    __syncthreads();
    if (tid == 0)
      for (int j = 0; j < nj; ++j) vel[j] = (par[j] >= 0 ? vel[par[j]] : zero6<T>()) + qdv[j] * col[j];
    __syncthreads();
    for (int j = tid; j < nj; j += nth) {
      dcol[j] = crossm(vel[j], col[j]);
      Ic[j] = mat_of(additive(world_inertia(JR[j], oR[j], op[j])));
      Bc[j] = coriolis_body(Ic[j], vel[j]);
    }
    __syncthreads();
    for (int k = tid; k < 72; k += nth)  // entry k of Ic (k < 36) or of Bc, leaves to root
      for (int j = nj - 1; j > 0; --j) {
        const int p = par[j];
        if (p < 0) continue;
        if (k < 36) Ic[p].a[k] += Ic[j].a[k];
        else Bc[p].a[k - 36] += Bc[j].a[k - 36];
      }
    __syncthreads();
    for (int j = tid; j < nj; j += nth) {
      F1[j] = mul6(Ic[j], dcol[j]) + mul6(Bc[j], col[j]);
      G[j] = mul6(Ic[j], col[j]);
      H[j] = mul6t(Bc[j], col[j]);
    }
    __syncthreads();
    T *Cb = Cout + b * nv * nv;
    for (int e = tid; e < nv * nv; e += nth) {
      const int r = joint_of_dof[e / nv], c = joint_of_dof[e % nv];
      T v = T(0);
      if (r <= c && c < r + JR[r].subtree) v = dot6(col[r], F1[c]);  // r ancestor-or-self of c
      else if (c < r && r < c + JR[c].subtree) v = dot6(G[r], dcol[c]) + dot6(H[r], col[c]);  // c strict ancestor
      Cb[e] = v;
    }
    return;
  }
  for (int j = 0; j < nj; ++j) {
    const JointRec<T> &J = JR[j];
    int p = J.parent, u = J.vuser;
    T qj = qb[u], s, c;
    sincos_t(qj, &s, &c);
    joint_world(J, qj, s, c, p >= 0 ? oR[p] : nullptr, p >= 0 ? op[p] : V3<T>{0, 0, 0}, p < 0, oR[j], op[j]);
    col[j] = jcol(J, oR[j], op[j]);
    vel[j] = (p >= 0 ? vel[p] : zero6<T>()) + qd[b * nv + u] * col[j];
    dcol[j] = crossm(vel[j], col[j]);
    Ic[j] = mat_of(additive(world_inertia(J, oR[j], op[j])));
    Bc[j] = coriolis_body(Ic[j], vel[j]);
  }
  for (int j = nj - 1; j > 0; --j) {
    int p = JR[j].parent;
    if (p >= 0)
      for (int k = 0; k < 36; ++k) { Ic[p].a[k] += Ic[j].a[k]; Bc[p].a[k] += Bc[j].a[k]; }
  }
  T *Cb = Cout + b * nv * nv;
  for (int k = 0; k < nv * nv; ++k) Cb[k] = T(0);
  for (int j = 0; j < nj; ++j) {
    int uj = JR[j].vuser;
    // k ancestor-or-self of j: C_kj = col_k . (Ic_j dcol_j + Bc_j col_j)
    S6<T> F1 = mul6(Ic[j], dcol[j]) + mul6(Bc[j], col[j]);
    for (int k = j; k >= 0; k = JR[k].parent) Cb[JR[k].vuser * nv + uj] = dot6(col[k], F1);
    // i strict ancestor of j: C_ji = (Ic_j col_j) . dcol_i + (Bc_j^T col_j) . col_i
    S6<T> G = mul6(Ic[j], col[j]), H = mul6t(Bc[j], col[j]);
    for (int i = JR[j].parent; i >= 0; i = JR[i].parent) Cb[uj * nv + JR[i].vuser] = dot6(G, dcol[i]) + dot6(H, col[i]);
  }
}

// One CLIK problem (pinocchio_model.cpp:228-300) of a batch sharing the target link.
template <class T, int MJ>
__device__ void clik_problem(const ModelTable<T> *__restrict__ M, const ClikParams &P, const T *__restrict__ target,
                             const T *__restrict__ q0, T *__restrict__ q_out, int32_t *__restrict__ success,
                             T *__restrict__ err_out, int32_t *__restrict__ iters, int64_t prob) {
  const int nv = M->nv, nc = P.nchain;
  const JointRec<T> *JR = joints_of(M);
  const LinkRec<T> &Lk = links_of(M)[P.link];
  const int64_t tgt = P.ntargets ? prob % P.ntargets : prob;
  const int32_t start = P.ntargets ? (int32_t)(prob / P.ntargets) : 0;
  if (P.solved && __hip_atomic_load(P.solved + tgt, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < start) {
    // a lower start already converged for this target: skip
    T *qo = q_out + prob * nv;
    for (int u = 0; u < nv; ++u) qo[u] = q0 ? q0[prob * nv + u] : T(0);
    success[prob] = 0;
    for (int r = 0; r < 6; ++r) err_out[prob * 6 + r] = T(1e30);
    if (iters) iters[prob] = -1;
    return;
  }
  {
    // desired pose of the link's joint frame: oMdes = target * link2joint^-1
    T Rt[9], Rd[9];
    V3<T> pt;
    pose_to_mat(target + tgt * 7, Rt, pt);
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j)
        Rd[3 * i + j] = Rt[3 * i] * Lk.R[3 * j] + Rt[3 * i + 1] * Lk.R[3 * j + 1] + Rt[3 * i + 2] * Lk.R[3 * j + 2];
    V3<T> pd = pt - mv(Rd, load3(Lk.p));
    T qc[MJ], qbest[MJ], msk[MJ];
    for (int k = 0; k < nc; ++k) {
      int u = JR[P.chain[k]].vuser;
      qc[k] = q0 ? q0[prob * nv + u] : T(0);
      qbest[k] = qc[k];
      msk[k] = P.mask ? T(P.mask[prob * P.mask_ld + u]) : T(1);
    }
    T minErr = T(1e10), bestErr[6] = {0, 0, 0, 0, 0, 0};
    bool ok = false;
    int it = 0;
    for (;; ++it) {
      T oR[MJ][9];
      V3<T> op[MJ];
      T Rj[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
      V3<T> pj = {0, 0, 0};
      for (int k = 0; k < nc; ++k) {
        const JointRec<T> &J = JR[P.chain[k]];
        T s, c;
        sincos_t(qc[k], &s, &c);
        joint_world(J, qc[k], s, c, k > 0 ? oR[k - 1] : nullptr, k > 0 ? op[k - 1] : V3<T>{0, 0, 0}, k == 0, oR[k],
                    op[k]);
      }
      if (nc > 0) {
        for (int k = 0; k < 9; ++k) Rj[k] = oR[nc - 1][k];
        pj = op[nc - 1];
      }
      // iMd = oMj^-1 oMdes
      T Ri[9];
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) Ri[3 * i + j] = Rj[i] * Rd[j] + Rj[3 + i] * Rd[3 + j] + Rj[6 + i] * Rd[6 + j];
      V3<T> pi = mtv(Rj, pd - pj);
      S6<T> e = log6_dev(Ri, pi);
      T ev[6];
      vec_of(e, ev);
      T en = sqrt(dot6(e, e));
      if (en < minErr) {
        minErr = en;
        for (int k = 0; k < nc; ++k) qbest[k] = qc[k];
        for (int r = 0; r < 6; ++r) bestErr[r] = ev[r];
      }
      if (en < T(P.eps)) { ok = true; break; }
      if (it >= P.max_iter) break;
      // Jlog6 of iMd^-1 = (Ri^T, -Ri^T pi)
      T Rinv[9];
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) Rinv[3 * i + j] = Ri[3 * j + i];
      T Jl[36];
      jlog6_dev(Rinv, T(-1) * mtv(Ri, pi), Jl);
      // J J^T + damp I and its LDL^T solve in FP64 whatever T is: with
      // damp = 1e-6 the 6x6 system near a singular pose has cond ~1e6+.
      T Jk[MJ][6];
      double JJt[36];
      for (int k = 0; k < 36; ++k) JJt[k] = 0.0;
      for (int k = 0; k < nc; ++k) {
        S6<T> cw = jcol(JR[P.chain[k]], oR[k], op[k]);
        S6<T> cl = to_local(Rj, pj, cw);
        T cv[6];
        vec_of(cl, cv);
        for (int r = 0; r < 6; ++r) {
          T acc = T(0);
          for (int m = 0; m < 6; ++m) acc += Jl[r * 6 + m] * cv[m];
          Jk[k][r] = -acc * msk[k];
        }
        for (int r = 0; r < 6; ++r)
          for (int m = 0; m <= r; ++m) JJt[r * 6 + m] += double(Jk[k][r]) * double(Jk[k][m]);
      }
      for (int r = 0; r < 6; ++r) {
        JJt[r * 6 + r] += P.damp;
        for (int m = 0; m < r; ++m) JJt[m * 6 + r] = JJt[r * 6 + m];
      }
      double y[6];
      for (int r = 0; r < 6; ++r) y[r] = double(ev[r]);
      ldlt6_solve(JJt, y);
      for (int k = 0; k < nc; ++k) {
        T v = T(0);
        for (int r = 0; r < 6; ++r) v -= Jk[k][r] * T(y[r]);
        qc[k] += v * T(P.dt);
      }
    }
    // result: chain dofs from the best iterate, the rest from q0; every
    // continuous dof wrapped by atan2 (posP2S over the whole configuration,
    // pinocchio_model.cpp:87,299)
    T *qo = q_out + prob * nv;
    for (int j = 0; j < M->nj; ++j) {
      const JointRec<T> &J = JR[j];
      T v = q0 ? q0[prob * nv + J.vuser] : T(0);
      if (J.kind == KIND_CONTINUOUS) {
        T s, c;
        sincos_t(v, &s, &c);
        v = atan2(s, c);
      }
      qo[J.vuser] = v;
    }
    for (int k = 0; k < nc; ++k) {
      const JointRec<T> &J = JR[P.chain[k]];
      T v = qbest[k];
      if (J.kind == KIND_CONTINUOUS) {
        T s, c;
        sincos_t(v, &s, &c);
        v = atan2(s, c);
      }
      qo[J.vuser] = v;
    }
    success[prob] = ok ? 1 : 0;
    for (int r = 0; r < 6; ++r) err_out[prob * 6 + r] = bestErr[r];
    if (iters) iters[prob] = it;
    if (ok && P.solved) __hip_atomic_fetch_min(P.solved + tgt, start, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}

template <class T, int MJ>
__global__ __launch_bounds__(64) void k_clik(const ModelTable<T> *__restrict__ M, int64_t B, ClikParams P,
                                             const T *__restrict__ target, const T *__restrict__ q0,
                                             T *__restrict__ q_out, int32_t *__restrict__ success,
                                             T *__restrict__ err_out, int32_t *__restrict__ iters) {
  if (!P.persistent_waves) {
    const int64_t prob = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (prob < B) clik_problem<T, MJ>(M, P, target, q0, q_out, success, err_out, iters, prob);
    return;
  }
  // persistent: every lane pulls problems from the queue until it runs dry
  for (int64_t prob = atomicAdd(P.counter, 1); prob < B; prob = atomicAdd(P.counter, 1))
    clik_problem<T, MJ>(M, P, target, q0, q_out, success, err_out, iters, prob);
}

// One CLIK problem per block of 64 threads (one problem, the per-call path):
// the chain joints' own transforms, Jacobian columns and updates one joint per
// thread and the 21 entries of J J^T one per thread; the chain composition,
// the SE(3) log with Jlog6 of its inverse (from one rotation log, the same
// values as log6_dev and jlog6_dev) and the 6x6 solve on thread 0. Each
// quantity has clik_problem's formula and summation order; the iterates agree
// with it to rounding (measured: converged configurations within 1e-10).
template <class T, int MJ>
__global__ __launch_bounds__(64) void k_clik_block(const ModelTable<T> *__restrict__ M, ClikParams P,
                                                   const T *__restrict__ target, const T *__restrict__ q0,
                                                   T *__restrict__ q_out, int32_t *__restrict__ success,
                                                   T *__restrict__ err_out, int32_t *__restrict__ iters) {
  const int tid = threadIdx.x, nth = blockDim.x;
  const int nj = M->nj, nc = P.nchain;
  const JointRec<T> *JR = joints_of(M);
  __shared__ T qc[MJ], qbest[MJ], msk[MJ], Jk[MJ][6], oR[MJ][9];
  __shared__ V3<T> op[MJ];
  __shared__ int in_chain[MJ];
  __shared__ T Rd[9], Rj[9], Jl[36], ev[6];
  __shared__ V3<T> pd, pj;
  __shared__ double JJt[36], y[6];
  __shared__ int state;  // bit 0: stop, bit 1: best iterate so far
  if (tid == 0) {
    // desired pose of the link's joint frame: oMdes = target * link2joint^-1
    const LinkRec<T> &Lk = links_of(M)[P.link];
    T Rt[9];
    V3<T> pt;
    pose_to_mat(target, Rt, pt);
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j)
        Rd[3 * i + j] = Rt[3 * i] * Lk.R[3 * j] + Rt[3 * i + 1] * Lk.R[3 * j + 1] + Rt[3 * i + 2] * Lk.R[3 * j + 2];
    pd = pt - mv(Rd, load3(Lk.p));
  }
  for (int j = tid; j < nj; j += nth) in_chain[j] = 0;
  __syncthreads();
  for (int k = tid; k < nc; k += nth) {
    const int u = JR[P.chain[k]].vuser;
    qc[k] = q0 ? q0[u] : T(0);
    qbest[k] = qc[k];
    msk[k] = P.mask ? T(P.mask[u]) : T(1);
    in_chain[P.chain[k]] = 1;
  }
  T minErr = T(1e10), bestErr[6] = {0, 0, 0, 0, 0, 0};  // thread 0's
  bool ok = false;
  int it = 0;
  for (;; ++it) {
    for (int k = tid; k < nc; k += nth) {  // the joints' own transforms
      T s, c;
      sincos_t(qc[k], &s, &c);
      joint_world(JR[P.chain[k]], qc[k], s, c, static_cast<const T *>(nullptr), V3<T>{0, 0, 0}, true, oR[k], op[k]);
    }
    __syncthreads();
    if (tid == 0) {
      for (int k = 1; k < nc; ++k) {  // root to link along the chain
        T R[9];
        mm(oR[k - 1], oR[k], R);
        op[k] = mv(oR[k - 1], op[k]) + op[k - 1];
        for (int i = 0; i < 9; ++i) oR[k][i] = R[i];
      }
      for (int i = 0; i < 9; ++i) Rj[i] = nc > 0 ? oR[nc - 1][i] : T(i % 4 == 0);
      pj = nc > 0 ? op[nc - 1] : V3<T>{0, 0, 0};
      // iMd = oMj^-1 oMdes
      T Ri[9];
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) Ri[3 * i + j] = Rj[i] * Rd[j] + Rj[3 + i] * Rd[3 + j] + Rj[6 + i] * Rd[6 + j];
      const V3<T> pi = mtv(Rj, pd - pj);
      const S6<T> e = log6_jlog6_inv(Ri, pi, Jl);
      T evl[6];
      vec_of(e, evl);
      const T en = sqrt(dot6(e, e));
      int st = 0;
      if (en < minErr) {
        minErr = en;
        st |= 2;
        for (int r = 0; r < 6; ++r) bestErr[r] = evl[r];
      }
      if (en < T(P.eps)) {
        ok = true;
        st |= 1;
      } else if (it >= P.max_iter) {
        st |= 1;
      }
      for (int r = 0; r < 6; ++r) ev[r] = evl[r];
      state = st;
    }
    __syncthreads();
    const int st = state;
    if (st & 2)
      for (int k = tid; k < nc; k += nth) qbest[k] = qc[k];
    if (st & 1) break;
    for (int k = tid; k < nc; k += nth) {  // rows of J (chain joint k) in the frame of the error
      const S6<T> cw = jcol(JR[P.chain[k]], oR[k], op[k]);
      const S6<T> cl = to_local(Rj, pj, cw);
      T cv[6];
      vec_of(cl, cv);
      for (int r = 0; r < 6; ++r) {
        T acc = T(0);
        for (int m = 0; m < 6; ++m) acc += Jl[r * 6 + m] * cv[m];
        Jk[k][r] = -acc * msk[k];
      }
    }
    __syncthreads();
    for (int ent = tid; ent < 21; ent += nth) {  // lower triangle of J J^T, FP64
      int r = 0;
      while ((r + 1) * (r + 2) / 2 <= ent) ++r;
      const int m = ent - r * (r + 1) / 2;
      double acc = 0.0;
      for (int k = 0; k < nc; ++k) acc += double(Jk[k][r]) * double(Jk[k][m]);
      JJt[r * 6 + m] = acc;
    }
    __syncthreads();
    if (tid == 0) {
      for (int r = 0; r < 6; ++r) {
        JJt[r * 6 + r] += P.damp;
        for (int m = 0; m < r; ++m) JJt[m * 6 + r] = JJt[r * 6 + m];
      }
      double yl[6], A[36];
      for (int r = 0; r < 6; ++r) yl[r] = double(ev[r]);
      for (int k = 0; k < 36; ++k) A[k] = JJt[k];
      ldlt6_solve(A, yl);
      for (int r = 0; r < 6; ++r) y[r] = yl[r];
    }
    __syncthreads();
    for (int k = tid; k < nc; k += nth) {
      T v = T(0);
      for (int r = 0; r < 6; ++r) v -= Jk[k][r] * T(y[r]);
      qc[k] += v * T(P.dt);
    }
    __syncthreads();
  }
  // result: chain dofs from the best iterate, the rest from q0; every
  // continuous dof wrapped by atan2 (posP2S over the whole configuration,
  // pinocchio_model.cpp:87,299)
  for (int j = tid; j < nj; j += nth) {
    if (in_chain[j]) continue;
    const JointRec<T> &J = JR[j];
    T v = q0 ? q0[J.vuser] : T(0);
    if (J.kind == KIND_CONTINUOUS) {
      T s, c;
      sincos_t(v, &s, &c);
      v = atan2(s, c);
    }
    q_out[J.vuser] = v;
  }
  for (int k = tid; k < nc; k += nth) {
    const JointRec<T> &J = JR[P.chain[k]];
    T v = qbest[k];
    if (J.kind == KIND_CONTINUOUS) {
      T s, c;
      sincos_t(v, &s, &c);
      v = atan2(s, c);
    }
    q_out[J.vuser] = v;
  }
  if (tid == 0) {
    success[0] = ok ? 1 : 0;
    for (int r = 0; r < 6; ++r) err_out[r] = bestErr[r];
    if (iters) iters[0] = it;
  }
}

// Problems of several models (segments) in one persistent launch: a lane takes
// the next global problem index and runs it on the segment that owns it.
template <int MJ> __global__ __launch_bounds__(64) void k_clik_hetero(HeteroClikArgs A) {
  for (int64_t g = atomicAdd(A.counter, 1); g < A.total; g = atomicAdd(A.counter, 1)) {
    int s = 0;
    while (s + 1 < A.nseg && g >= A.seg[s + 1].first) ++s;
    const ClikSegment &S = A.seg[s];
    clik_problem<float, MJ>(S.tab, S.P, S.target, S.q0, S.q_out, S.success, S.err, S.iters, g - S.first);
  }
}

__global__ void k_fill_i32(int32_t *dst, int32_t value, int64_t n) {
  int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) dst[i] = value;
}

// Per-joint arrays are sized by the joint-count class MJ; the private segment
// (scratch) of a thread-per-environment kernel grows with MJ, so the classes
// are kept tight.
inline int pick_mj(int nj) {
  return nj <= 8 ? 8 : nj <= 16 ? 16 : nj <= 24 ? 24 : nj <= 32 ? 32 : nj <= 48 ? 48 : 64;
}

#define SAPAMD_DISPATCH_MJ(nj, KERNEL, ...)                                                      \
  switch (pick_mj(nj)) {                                                                         \
    case 8: hipLaunchKernelGGL((KERNEL<T, 8>), __VA_ARGS__); break;                             \
    case 16: hipLaunchKernelGGL((KERNEL<T, 16>), __VA_ARGS__); break;                           \
    case 24: hipLaunchKernelGGL((KERNEL<T, 24>), __VA_ARGS__); break;                           \
    case 32: hipLaunchKernelGGL((KERNEL<T, 32>), __VA_ARGS__); break;                           \
    case 48: hipLaunchKernelGGL((KERNEL<T, 48>), __VA_ARGS__); break;                           \
    default: hipLaunchKernelGGL((KERNEL<T, 64>), __VA_ARGS__); break;                           \
  }
// Kernels with per-joint arrays: one environment -> LDS arrays, one block of 64
// threads (kernels without a shared-array path run it on thread 0).
#define SAPAMD_DISPATCH_MJ_ARR(nj, B, KERNEL, s, ...)                                                     \
  if ((B) == 1) {                                                                                          \
    switch (pick_mj(nj)) {                                                                                 \
      case 8: hipLaunchKernelGGL((KERNEL<T, 8, true>), dim3(1), dim3(64), 0, s, __VA_ARGS__); break;       \
      case 16: hipLaunchKernelGGL((KERNEL<T, 16, true>), dim3(1), dim3(64), 0, s, __VA_ARGS__); break;     \
      case 24: hipLaunchKernelGGL((KERNEL<T, 24, true>), dim3(1), dim3(64), 0, s, __VA_ARGS__); break;     \
      case 32: hipLaunchKernelGGL((KERNEL<T, 32, true>), dim3(1), dim3(64), 0, s, __VA_ARGS__); break;     \
      case 48: hipLaunchKernelGGL((KERNEL<T, 48, true>), dim3(1), dim3(64), 0, s, __VA_ARGS__); break;     \
      default: hipLaunchKernelGGL((KERNEL<T, 64, true>), dim3(1), dim3(64), 0, s, __VA_ARGS__); break;     \
    }                                                                                                      \
  } else {                                                                                                 \
    const dim3 g_ = grid_for(B);                                                                           \
    switch (pick_mj(nj)) {                                                                                 \
      case 8: hipLaunchKernelGGL((KERNEL<T, 8, false>), g_, dim3(64), 0, s, __VA_ARGS__); break;          \
      case 16: hipLaunchKernelGGL((KERNEL<T, 16, false>), g_, dim3(64), 0, s, __VA_ARGS__); break;        \
      case 24: hipLaunchKernelGGL((KERNEL<T, 24, false>), g_, dim3(64), 0, s, __VA_ARGS__); break;        \
      case 32: hipLaunchKernelGGL((KERNEL<T, 32, false>), g_, dim3(64), 0, s, __VA_ARGS__); break;        \
      case 48: hipLaunchKernelGGL((KERNEL<T, 48, false>), g_, dim3(64), 0, s, __VA_ARGS__); break;        \
      default: hipLaunchKernelGGL((KERNEL<T, 64, false>), g_, dim3(64), 0, s, __VA_ARGS__); break;        \
    }                                                                                                      \
  }

inline dim3 grid_for(int64_t B) { return dim3((unsigned)((B + 63) / 64)); }

}  // namespace

namespace {
// Philox4x32-10 (Salmon, Moraes, Dror and Shaw, "Parallel random numbers: as
// easy as 1, 2, 3", SC 2011): ten rounds of two 32 x 32 -> 64-bit products,
// the key bumped by the Weyl constants between rounds.
__device__ __forceinline__ uint4 philox4x32_10(uint4 c, uint32_t k0, uint32_t k1) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    const uint64_t p0 = (uint64_t)0xD2511F53u * c.x, p1 = (uint64_t)0xCD9E8D57u * c.z;
    c = make_uint4((uint32_t)(p1 >> 32) ^ c.y ^ k0, (uint32_t)p1, (uint32_t)(p0 >> 32) ^ c.w ^ k1, (uint32_t)p0);
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  return c;
}

// Entry (b, u) of a batch of random configurations: uniform in [lo_u, hi_u)
// from the 53-bit fraction of Philox4x32-10 at counter (b, u) under the seed.
template <class T>
__global__ void k_random_qpos(int64_t B, int nv, uint64_t seed, RandomLimits lim, T *__restrict__ out) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= B * nv) return;
  const int64_t b = i / nv;
  const int u = (int)(i - b * nv);
  const uint4 r = philox4x32_10(make_uint4((uint32_t)b, (uint32_t)((uint64_t)b >> 32), (uint32_t)u, 0u),
                                (uint32_t)seed, (uint32_t)(seed >> 32));
  const double f = ((double)(r.x >> 5) * 67108864.0 + (double)(r.y >> 6)) * (1.0 / 9007199254740992.0);
  // separately rounded product and sum (no contraction into an FMA), so the
  // value is reproducible by any IEEE double evaluation of the same formula
  out[i] = (T)__dadd_rn(lim.lo[u], __dmul_rn(lim.hi[u] - lim.lo[u], f));
}
}  // namespace

template <class T>
hipError_t launch_random_qpos(int64_t B, int nv, uint64_t seed, const RandomLimits &lim, T *out, hipStream_t s) {
  if (B <= 0 || nv <= 0) return hipSuccess;
  if (nv > kRandomMaxDofs) return hipErrorInvalidValue;
  const int64_t n = B * nv;
  hipLaunchKernelGGL(k_random_qpos<T>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, B, nv, seed, lim, out);
  return hipGetLastError();
}
template hipError_t launch_random_qpos<float>(int64_t, int, uint64_t, const RandomLimits &, float *, hipStream_t);
template hipError_t launch_random_qpos<double>(int64_t, int, uint64_t, const RandomLimits &, double *, hipStream_t);

hipError_t launch_fill_i32(int32_t *dst, int32_t value, int64_t n, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_fill_i32, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, dst, value, n);
  return hipGetLastError();
}

template <class T>
hipError_t launch_kinematics(const ModelTable<T> *tab, int nj, int64_t B, int flags, const T *q, const T *qd,
                             const T *qdd, KinOut<T> out, hipStream_t s) {
  if (B <= 0) return hipSuccess;
  SAPAMD_DISPATCH_MJ_ARR(nj, B, k_kinematics, s, tab, B, flags, q, qd, qdd, out);
  return hipGetLastError();
}
template <class T>
hipError_t launch_link_local_jacobian(const ModelTable<T> *tab, int nj, int64_t B, int link, const T *q, T *J,
                                      hipStream_t s) {
  if (B <= 0) return hipSuccess;
  SAPAMD_DISPATCH_MJ_ARR(nj, B, k_link_local_jacobian, s, tab, B, link, q, J);
  return hipGetLastError();
}
template <class T>
hipError_t launch_crba(const ModelTable<T> *tab, int nj, int64_t B, const T *q, T *M, hipStream_t s) {
  if (B <= 0) return hipSuccess;
  SAPAMD_DISPATCH_MJ_ARR(nj, B, k_crba, s, tab, B, q, M);
  return hipGetLastError();
}
template <class T>
hipError_t launch_coriolis(const ModelTable<T> *tab, int nj, int64_t B, const T *q, const T *qd, T *C,
                           hipStream_t s) {
  if (B <= 0) return hipSuccess;
  SAPAMD_DISPATCH_MJ_ARR(nj, B, k_coriolis, s, tab, B, q, qd, C);
  return hipGetLastError();
}
template <class T>
hipError_t launch_aba(const ModelTable<T> *tab, int nj, int64_t B, const T *q, const T *qd, const T *tau, T *qdd,
                      hipStream_t s) {
  if (B <= 0) return hipSuccess;
  SAPAMD_DISPATCH_MJ_ARR(nj, B, k_aba, s, tab, B, q, qd, tau, qdd);
  return hipGetLastError();
}
template <class T>
hipError_t launch_clik(const ModelTable<T> *tab, int nj, int64_t B, const ClikParams &P, const T *target,
                       const T *q0, T *q_out, int32_t *success, T *err, int32_t *iters, hipStream_t s) {
  if (B <= 0) return hipSuccess;
  if (B == 1 && !P.persistent_waves && !P.ntargets && !P.solved && nj <= 64) {  // one problem: a block of 64 threads
    SAPAMD_DISPATCH_MJ(nj, k_clik_block, dim3(1), dim3(64), 0, s, tab, P, target, q0, q_out, success, err, iters);
    return hipGetLastError();
  }
  dim3 grid = P.persistent_waves ? dim3((unsigned)P.persistent_waves) : grid_for(B);
  SAPAMD_DISPATCH_MJ(P.nchain, k_clik, grid, dim3(64), 0, s, tab, B, P, target, q0, q_out, success, err, iters);
  return hipGetLastError();
}

hipError_t launch_clik_hetero(const HeteroClikArgs &A, int max_chain, int waves, hipStream_t s) {
  if (A.total <= 0) return hipSuccess;
  const dim3 grid((unsigned)waves);
  if (max_chain <= 8) hipLaunchKernelGGL((k_clik_hetero<8>), grid, dim3(64), 0, s, A);
  else if (max_chain <= 16) hipLaunchKernelGGL((k_clik_hetero<16>), grid, dim3(64), 0, s, A);
  else if (max_chain <= 32) hipLaunchKernelGGL((k_clik_hetero<32>), grid, dim3(64), 0, s, A);
  else hipLaunchKernelGGL((k_clik_hetero<64>), grid, dim3(64), 0, s, A);
  return hipGetLastError();
}

#define SAPAMD_INST(T)                                                                                           \
  template hipError_t launch_kinematics<T>(const ModelTable<T> *, int, int64_t, int, const T *, const T *,     \
                                           const T *, KinOut<T>, hipStream_t);                                 \
  template hipError_t launch_link_local_jacobian<T>(const ModelTable<T> *, int, int64_t, int, const T *, T *,  \
                                                    hipStream_t);                                              \
  template hipError_t launch_crba<T>(const ModelTable<T> *, int, int64_t, const T *, T *, hipStream_t);         \
  template hipError_t launch_coriolis<T>(const ModelTable<T> *, int, int64_t, const T *, const T *, T *,       \
                                         hipStream_t);                                                         \
  template hipError_t launch_aba<T>(const ModelTable<T> *, int, int64_t, const T *, const T *, const T *, T *,  \
                                    hipStream_t);                                                              \
  template hipError_t launch_clik<T>(const ModelTable<T> *, int, int64_t, const ClikParams &, const T *,        \
                                     const T *, T *, int32_t *, T *, int32_t *, hipStream_t);

SAPAMD_INST(float)
SAPAMD_INST(double)

}  // namespace sapamd
