// Lanes-per-joint kernels for kinematic trees (FP32), and the heterogeneous
// launch that runs several trees in one grid.
//
// An environment occupies G lanes (G = 8, 16 or 32 >= nj); lane j owns joint j
// of the depth-first (preorder) joint order. Two tree scans replace the serial
// sweeps:
//  * root-to-joint prefix (FK products, velocities, accelerations): pointer
//    jumping over the precomputed 2^k-th ancestors, x_j <- x_anc_k(j) (+) x_j
//    for k = 0 .. log2(depth), exchanging registers with ds_bpermute;
//  * subtree sums (forces, composite inertias): in preorder the subtree of j is
//    the lane range [j, j + size_j). Doubling builds block sums
//    B_k[j] = sum of 2^k lanes from j, and F_j adds the blocks of the binary
//    decomposition of size_j. Exact, without prefix differences (no cancellation).
// The CRBA lane j writes row j of M: for i an ancestor-or-self of j the entry is
// col_i . (Yc_j col_j), for i a descendant it is col_j . (Yc_i col_i), else 0.
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <algorithm>

#include "lanes.h"
#include "launch.h"
#include "spatial.h"
#include "tree.h"

namespace sapamd {

namespace {

using namespace lanes;

constexpr int kMaxTreeJoints = 32;
constexpr int kTreeTableWords4 = (int)((kHeaderBytes + sizeof(JointRec<float>) * kMaxTreeJoints) / 16);

struct TreeIO {
  const float *q, *qd, *qdd;
  float *poses, *J, *tau, *M;
  // row strides in elements (0 = contiguous: nv, L * 7 and 7)
  int64_t ld_q = 0, ld_qd = 0, ld_qdd = 0, ld_tau = 0, pose_env = 0, pose_link = 0;
};

// Dynamic LDS of a launch that computes M: per lane its joint column and
// composite-inertia force (12 floats).
__host__ __device__ constexpr int crba_lds_bytes(int) { return 256 * 12 * 4; }

// One environment per G-lane group; M is the LDS copy of the model table;
// crba is the block's dynamic LDS (used when flags has KF_CRBA): 3 float4 per
// lane.
template <int G>
__device__ __forceinline__ void tree_body(const ModelTable<float> *M, const ModelTable<float> *Mg, int64_t B,
                                          int64_t env, int flags, TreeIO io, float4 *crba) {
  const int g = threadIdx.x & (G - 1);
  const bool live = env < B;
  const int64_t e = live ? env : B - 1;
  const int nj = M->nj, nv = M->nv, L = M->nlinks;
  const bool dyn = flags & KF_RNEA;
  const bool valid = g < nj;
  const JointRec<float> &J = joints_of(M)[valid ? g : nj - 1];
  Tree<G> T;
  T.g = g;
  T.base4 = ((threadIdx.x & 63) & ~(G - 1)) * 4;
#pragma unroll
  for (int k = 0; k < Tree<G>::K; ++k) T.anc[k] = valid ? J.anc[k] : -1;
  T.size = valid ? J.subtree : 0;

  // dofs: lane g loads user dof g, then the joint -> dof permutation via bpermute
  const int gc = g < nv ? g : nv - 1;
  float qu = io.q[e * (io.ld_q ? io.ld_q : nv) + gc], qdu = io.qd[e * (io.ld_qd ? io.ld_qd : nv) + gc],
        qddu = io.qdd[e * (io.ld_qdd ? io.ld_qdd : nv) + gc];
  const int src = T.base4 + 4 * J.vuser;
  float qj = bperm(src, qu), qdj = bperm(src, qdu), qddj = bperm(src, qddu);
  if (!valid) qj = qdj = qddj = 0.f;
  if (!dyn || (flags & KF_QD0)) qdj = 0.f;
  if (!dyn || (flags & KF_QDD0)) qddj = 0.f;

  float s, c;
  joint_sincos(qj, &s, &c);
  Xf X;
  joint_world(J, qj, s, c, static_cast<const float *>(nullptr), V3<float>{0, 0, 0}, true, X.R, X.p);
  if (!valid) X = identity_xf();
  T.prefix(X);
  const S6<float> col = valid ? jcol(J, X.R, X.p) : zero6<float>();
  const int u = J.vuser;
  constexpr bool NT = G >= 8;  // streaming output stores (see st())
  float *const prow = io.poses + e * (io.pose_env ? io.pose_env : (int64_t)L * 7);
  const int pl = io.pose_link ? (int)io.pose_link : 7;

  // Jacobian columns and link poses first: their stores drain while the
  // dynamics compute
  if ((flags & KF_JFULL) && live && valid) {
    float *Jb = io.J + e * 6 * nv + u;
    st<NT>(col.l.x, Jb + 0 * nv); st<NT>(col.l.y, Jb + 1 * nv); st<NT>(col.l.z, Jb + 2 * nv);
    st<NT>(col.a.x, Jb + 3 * nv); st<NT>(col.a.y, Jb + 4 * nv); st<NT>(col.a.z, Jb + 5 * nv);
  }
  if ((flags & KF_POSES) && M->links_le1) {
    if (live && valid && J.link >= 0) {
      if (J.lident) {  // link frame = joint frame
        store_pose<NT>(prow + J.link * pl, X.R, X.p);
      } else {
        float Rw[9];
        mm(X.R, J.lR, Rw);
        store_pose<NT>(prow + J.link * pl, Rw, mv(X.R, load3(J.lp)) + X.p);
      }
    }
    if (live && g == 0 && M->root_link >= 0)
      store_pose_q<NT>(prow + M->root_link * pl, M->root_p, M->root_q);
  } else if (flags & KF_POSES) {
    for (int l0 = 0; l0 < L; l0 += G) {  // every lane takes part in the permutes
      const int l = l0 + g;
      const LinkRec<float> &LL = links_of(Mg)[l < L ? l : L - 1];
      const int jl = LL.joint < 0 ? 0 : LL.joint;
      const Xf W = bperm_xf(T.base4 + 4 * jl, X);
      float Rw[9];
      mm(W.R, LL.R, Rw);
      V3<float> pw = mv(W.R, load3(LL.p)) + W.p;
      if (LL.joint < 0) {
#pragma unroll
        for (int k = 0; k < 9; ++k) Rw[k] = LL.R[k];
        pw = load3(LL.p);
      }
      if (live && l < L) store_pose<NT>(prow + l * pl, Rw, pw);
    }
  }

  // composite inertia of the subtree (additive form about the world origin),
  // summed with the RNEA forces when both are computed
  const bool do_crba = flags & KF_CRBA;
  float y[10];
  if (do_crba) {
    const AInertia<float> Ya = additive(world_inertia_fast(J, X.R, X.p));
    const float yv[10] = {Ya.m, Ya.h.x, Ya.h.y, Ya.h.z, Ya.I[0], Ya.I[1], Ya.I[2], Ya.I[3], Ya.I[4], Ya.I[5]};
#pragma unroll
    for (int i = 0; i < 10; ++i) y[i] = valid ? yv[i] : 0.f;
  }
  if (dyn) {
    S6<float> v = qdj * col;
    T.prefix(v);
    S6<float> a = qddj * col + qdj * crossm(v, col);
    T.prefix(a);
    if (!(flags & KF_NOGRAV)) a.l = a.l - V3<float>{M->gravity[0], M->gravity[1], M->gravity[2]};
    const WInertia<float> Y = world_inertia_fast(J, X.R, X.p);
    S6<float> f = valid ? inertia_apply(Y, a) + crossf(v, inertia_apply(Y, v)) : zero6<float>();
    if (do_crba) {
      // one subtree pass for the forces and the composite inertias
      float x[16] = {f.l.x, f.l.y, f.l.z, f.a.x, f.a.y, f.a.z};
#pragma unroll
      for (int i = 0; i < 10; ++i) x[6 + i] = y[i];
      T.template subtree<16>(x);
      f = {{x[0], x[1], x[2]}, {x[3], x[4], x[5]}};
#pragma unroll
      for (int i = 0; i < 10; ++i) y[i] = x[6 + i];
    } else {
      T.subtree(f);
    }
    if (live && valid) st<NT>(dot6(col, f), io.tau + e * (io.ld_tau ? io.ld_tau : nv) + u);
  }
  if (do_crba) {
    if (!dyn) T.template subtree<10>(y);
    AInertia<float> Yc;
    Yc.m = y[0];
    Yc.h = {y[1], y[2], y[3]};
#pragma unroll
    for (int i = 0; i < 6; ++i) Yc.I[i] = y[4 + i];
    const S6<float> Fj = apply(Yc, col);
    // every lane publishes its column and force; lane j then reads those of
    // the other joints (one address across the group at a time) and stores
    // row u of M
    float4 *cf = crba + (threadIdx.x & ~(G - 1) & 255) * 3;
    cf[g * 3 + 0] = make_float4(col.l.x, col.l.y, col.l.z, col.a.x);
    cf[g * 3 + 1] = make_float4(col.a.y, col.a.z, Fj.l.x, Fj.l.y);
    cf[g * 3 + 2] = make_float4(Fj.l.z, Fj.a.x, Fj.a.y, Fj.a.z);
    wave_sync();
    // branch-free: every load is issued up front, both products computed and
    // selected. M is symmetric (both triangles come from the same dot6(S_a,
    // F_b) with the ancestor's column first), so lane u stores entry (i, u)
    // instead of (u, i): at each i the group writes one contiguous row.
    float *Mcol = io.M + e * nv * nv + u;
    const bool out = live && valid;
    // joint i's subtree size and user dof sit on lane i of every group (one
    // model per wavefront): read into scalars
#pragma unroll 4
    for (int i = 0; i < nj; ++i) {
      const float4 r0 = cf[i * 3 + 0], r1 = cf[i * 3 + 1], r2 = cf[i * 3 + 2];
      const S6<float> ci = {{r0.x, r0.y, r0.z}, {r0.w, r1.x, r1.y}};
      const S6<float> Fi = {{r1.z, r1.w, r2.x}, {r2.y, r2.z, r2.w}};
      const int sub_i = __builtin_amdgcn_readlane(T.size, i), vu_i = __builtin_amdgcn_readlane(u, i);
      const float dij = dot6(ci, Fj), dji = dot6(col, Fi);
      const bool i_anc_j = i <= g && g < i + sub_i;
      const bool j_anc_i = g < i && i < g + T.size;
      if (out) st<false>(i_anc_j ? dij : j_anc_i ? dji : 0.f, Mcol + (int64_t)vu_i * nv);
    }
  }
}

// Forward recomputation + reverse pass (formulas in kernels_adjoint.hip).
// Pose gradients need one user link per joint (links_le1).
template <int G>
__device__ __forceinline__ void tree_vjp_body(const ModelTable<float> *M, int64_t B, int64_t env, const float *q,
                                              const float *qd, const float *qdd, const float *gsrc,
                                              const VjpArgs<float> &A) {
  const int g = threadIdx.x & (G - 1);
  const bool live = env < B;
  const int64_t e = live ? env : B - 1;
  const int nj = M->nj, nv = M->nv, L = M->nlinks;
  const bool dyn = A.gtau || A.tau_star;
  const bool valid = g < nj;
  const JointRec<float> &J = joints_of(M)[valid ? g : nj - 1];
  Tree<G> T;
  T.g = g;
  T.base4 = ((threadIdx.x & 63) & ~(G - 1)) * 4;
#pragma unroll
  for (int k = 0; k < Tree<G>::K; ++k) T.anc[k] = valid ? J.anc[k] : -1;
  T.size = valid ? J.subtree : 0;

  const int64_t qi = e * nv + (g < nv ? g : nv - 1);
  const float qu = q[qi], qdu = qd[qi], qddu = qdd[qi], gu = gsrc[qi];
  V3<float> ptarget = {0.f, 0.f, 0.f};
  if (A.p_target) ptarget = {A.p_target[e * 3], A.p_target[e * 3 + 1], A.p_target[e * 3 + 2]};
  const int src = T.base4 + 4 * J.vuser;
  float qj = bperm(src, qu), qdj = bperm(src, qdu), qddj = bperm(src, qddu), gj = bperm(src, gu);
  if (!valid) qj = qdj = qddj = gj = 0.f;
  if (!dyn) qdj = qddj = gj = 0.f;

  float s, c;
  joint_sincos(qj, &s, &c);
  Xf X;
  joint_world(J, qj, s, c, static_cast<const float *>(nullptr), V3<float>{0, 0, 0}, true, X.R, X.p);
  if (!valid) X = identity_xf();
  T.prefix(X);
  const S6<float> col = valid ? jcol(J, X.R, X.p) : zero6<float>();
  const int u = J.vuser;

  float loss = 0.f;
  S6<float> FP = zero6<float>();
  if (valid && J.link >= 0) {
    const int l = J.link;
    float Rw[9];
    mm(X.R, J.lR, Rw);
    const V3<float> pw = mv(X.R, load3(J.lp)) + X.p;
    if (A.gposes) {
      float quat[4];
      quat_branchless(Rw, quat);
      const float *gp7 = A.gposes + (e * L + l) * 7;
      const V3<float> gp = {gp7[0], gp7[1], gp7[2]}, gq = {gp7[4], gp7[5], gp7[6]};
      const V3<float> uq = {quat[1], quat[2], quat[3]};
      FP = {gp, cross(pw, gp) + 0.5f * (quat[0] * gq - gp7[3] * uq + cross(uq, gq))};
    } else if (A.p_target && l == A.ee_link) {
      const V3<float> d = pw - ptarget;
      loss += dot(d, d);
      const V3<float> gp = 2.f * d;
      FP = {gp, cross(pw, gp)};
    }
  }
  S6<float> Hw = zero6<float>(), QS = zero6<float>(), v = zero6<float>(), a = zero6<float>(), F = zero6<float>(),
            w = zero6<float>();
  if (dyn) {
    v = qdj * col;
    T.prefix(v);
    a = qddj * col + qdj * crossm(v, col);
    T.prefix(a);
    a.l = a.l - V3<float>{M->gravity[0], M->gravity[1], M->gravity[2]};
    const WInertia<float> Y = world_inertia_fast(J, X.R, X.p);
    F = valid ? inertia_apply(Y, a) + crossf(v, inertia_apply(Y, v)) : zero6<float>();
    T.subtree(F);
    const float tj = dot6(col, F);
    if (A.tau && live && valid) A.tau[e * nv + u] = tj;
    if (A.tau_star) {
      const float d = valid ? tj - gj : 0.f;
      loss += d * d;
      gj = 2.f * d;
    }
    w = gj * col;
    T.prefix(w);
    const S6<float> Iw = inertia_apply(Y, w), Iv = inertia_apply(Y, v);
    if (valid) {
      Hw = Iw;
      QS = crossf(w, Iv) - crossf(v, Iw) + inertia_apply(Y, crossm(v, w));
    }
  }
  float acc[18] = {Hw.l.x, Hw.l.y, Hw.l.z, Hw.a.x, Hw.a.y, Hw.a.z, QS.l.x, QS.l.y, QS.l.z,
                   QS.a.x, QS.a.y, QS.a.z, FP.l.x, FP.l.y, FP.l.z, FP.a.x, FP.a.y, FP.a.z};
  T.template subtree<18>(acc);
  Hw = {{acc[0], acc[1], acc[2]}, {acc[3], acc[4], acc[5]}};
  QS = {{acc[6], acc[7], acc[8]}, {acc[9], acc[10], acc[11]}};
  FP = {{acc[12], acc[13], acc[14]}, {acc[15], acc[16], acc[17]}};
  if (live && valid) {
    float dq = dot6(col, FP);
    if (dyn) {
      const S6<float> wl = w - gj * col, vl = v - qdj * col;
      const S6<float> al = a - qddj * col - qdj * crossm(v, col);
      dq += dot6(col, crossf(vl, QS + crossf(vl, Hw)) - crossf(wl, F) - crossf(al, Hw));
      if (A.dqd) A.dqd[e * nv + u] = -dot6(col, QS + 2.f * crossf(vl, Hw));
      if (A.dqdd) A.dqdd[e * nv + u] = dot6(col, Hw);
    }
    if (A.dq) A.dq[e * nv + u] = dq;
  }
  if ((A.tau_star || A.p_target) && A.loss) {
    loss = group_sum<G>(g, T.base4, loss);
    if (live && g == 0) A.loss[e] = loss;
  }
}

__device__ __forceinline__ void stage_table(float4 *s_tab, const ModelTable<float> *Mg) {
  const int n4 = (int)((kHeaderBytes + sizeof(JointRec<float>) * Mg->nj) / 16);
  const float4 *src = reinterpret_cast<const float4 *>(Mg);
  for (int i = threadIdx.x; i < n4; i += blockDim.x) s_tab[i] = src[i];
}

template <int G>
__global__ __launch_bounds__(256) void k_tree_fused(const ModelTable<float> *__restrict__ Mg, int64_t B, int flags,
                                                    TreeIO io) {
  __shared__ float4 s_tab[kTreeTableWords4];
  extern __shared__ float4 s_crba[];  // crba_lds_bytes(G) when KF_CRBA, else empty
  stage_table(s_tab, Mg);
  __syncthreads();
  const int64_t env = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) / G;
  tree_body<G>(reinterpret_cast<const ModelTable<float> *>(s_tab), Mg, B, env, flags, io, s_crba);
}

template <int G>
__global__ __launch_bounds__(256) void k_tree_vjp(const ModelTable<float> *__restrict__ Mg, int64_t B,
                                                  const float *__restrict__ q, const float *__restrict__ qd,
                                                  const float *__restrict__ qdd, const float *__restrict__ gsrc,
                                                  VjpArgs<float> A) {
  __shared__ float4 s_tab[kTreeTableWords4];
  stage_table(s_tab, Mg);
  __syncthreads();
  const int64_t env = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) / G;
  tree_vjp_body<G>(reinterpret_cast<const ModelTable<float> *>(s_tab), B, env, q, qd, qdd, gsrc, A);
}

// Link poses and velocities written into a GPU simulator's body-state rows
// (13 floats: p xyz | q wxyz | v xyz | w xyz, world frame). Row 0 of an
// environment holds user link 0 (the simulator's root link): its pose and
// velocity are read first and place the articulation. With (R_m, p_m) a link's
// pose and (w_m, v_m) its angular and origin velocity in the model frame
// (from q, qd), and (R_0, p_0, w_0) read from row 0:
//   R_b = R_0 R_m0^T, p_b = p_0 - R_b p_m0          (the model frame in the world)
//   w_b = w_0 - R_b w_m0, u_b = v_0 - R_b v_m0      (its motion; u_b at point p_0)
//   pose_l = (R_b R_ml, R_b p_ml + p_b),  w_l = w_b + R_b w_ml,
//   v_l = u_b + w_b x (p_l - p_0) + R_b v_ml + w_l x (R_l c_l),
// v_0 being row 0's origin velocity (its stored centre-of-mass velocity minus
// w_0 x R_0 c_0) and the last term moving each origin velocity to the link's
// centre of mass c_l. Without qd only the poses (first 7 floats) are written.
struct LinkStateIO {
  const float *q, *qd;
  int64_t ld_q, ld_qd;
  float *body;
  int64_t env_stride, link_stride;
};

// Model-frame pose and velocity (angular, at the link origin) of user link l
// from the group's joint transforms X and joint velocities v (every lane calls
// this with its own l: the permutes run in all lanes).
template <int G>
__device__ __forceinline__ void link_model_state(const Tree<G> &T, const LinkRec<float> &LL, const Xf &X,
                                                 const S6<float> &v, float *Rm, V3<float> &pm, V3<float> &wm,
                                                 V3<float> &vm) {
  const int jl = LL.joint < 0 ? 0 : LL.joint;
  const Xf W = bperm_xf(T.base4 + 4 * jl, X);
  const S6<float> Vj = bperm6(T.base4 + 4 * jl, v);
  mm(W.R, LL.R, Rm);
  pm = mv(W.R, load3(LL.p)) + W.p;
  wm = Vj.a;
  vm = Vj.l + cross(Vj.a, pm);
  if (LL.joint < 0) {  // on the universe: fixed in the model frame
#pragma unroll
    for (int k = 0; k < 9; ++k) Rm[k] = LL.R[k];
    pm = load3(LL.p);
    wm = vm = V3<float>{0.f, 0.f, 0.f};
  }
}

template <int G>
__global__ __launch_bounds__(256) void k_link_state(const ModelTable<float> *__restrict__ Mg, int64_t B,
                                                    LinkStateIO io) {
  __shared__ float4 s_tab[kTreeTableWords4];
  stage_table(s_tab, Mg);
  __syncthreads();
  const ModelTable<float> *M = reinterpret_cast<const ModelTable<float> *>(s_tab);
  const int64_t env = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) / G;
  const int g = threadIdx.x & (G - 1);
  const bool live = env < B;
  const int64_t e = live ? env : B - 1;
  const int nj = M->nj, nv = M->nv, L = M->nlinks;
  const bool vel = io.qd != nullptr;
  const bool valid = g < nj;
  const JointRec<float> &J = joints_of(M)[valid ? g : nj - 1];
  Tree<G> T;
  T.g = g;
  T.base4 = ((threadIdx.x & 63) & ~(G - 1)) * 4;
#pragma unroll
  for (int k = 0; k < Tree<G>::K; ++k) T.anc[k] = valid ? J.anc[k] : -1;
  T.size = valid ? J.subtree : 0;

  const int gc = g < nv ? g : nv - 1;
  const float qu = io.q[e * io.ld_q + gc], qdu = vel ? io.qd[e * io.ld_qd + gc] : 0.f;
  const int src = T.base4 + 4 * J.vuser;
  float qj = bperm(src, qu), qdj = bperm(src, qdu);
  if (!valid) qj = qdj = 0.f;
  float s, c;
  joint_sincos(qj, &s, &c);
  Xf X;
  joint_world(J, qj, s, c, static_cast<const float *>(nullptr), V3<float>{0, 0, 0}, true, X.R, X.p);
  if (!valid) X = identity_xf();
  T.prefix(X);
  S6<float> v = zero6<float>();
  if (vel) {
    v = valid ? qdj * jcol(J, X.R, X.p) : zero6<float>();
    T.prefix(v);  // spatial velocity of each joint body, model frame, linear part at the model origin
  }

  // row 0: user link 0's world pose and velocity (read before any row is written)
  float *row = io.body + e * io.env_stride;
  const V3<float> p0 = {row[0], row[1], row[2]};
  float qw = row[3], qx = row[4], qy = row[5], qz = row[6];
  const float rn = rsqrtf(qw * qw + qx * qx + qy * qy + qz * qz);
  qw *= rn; qx *= rn; qy *= rn; qz *= rn;
  const float R0[9] = {1.f - 2.f * (qy * qy + qz * qz), 2.f * (qx * qy - qw * qz), 2.f * (qx * qz + qw * qy),
                       2.f * (qx * qy + qw * qz), 1.f - 2.f * (qx * qx + qz * qz), 2.f * (qy * qz - qw * qx),
                       2.f * (qx * qz - qw * qy), 2.f * (qy * qz + qw * qx), 1.f - 2.f * (qx * qx + qy * qy)};
  V3<float> vc0 = {0.f, 0.f, 0.f}, w0 = {0.f, 0.f, 0.f};
  if (vel) {
    vc0 = {row[7], row[8], row[9]};
    w0 = {row[10], row[11], row[12]};
  }
  const LinkRec<float> &L0 = links_of(Mg)[0];
  float Rm0[9];
  V3<float> pm0, wm0, vm0;
  link_model_state<G>(T, L0, X, v, Rm0, pm0, wm0, vm0);
  float Rb[9];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int k = 0; k < 3; ++k)
      Rb[3 * i + k] = R0[3 * i] * Rm0[3 * k] + R0[3 * i + 1] * Rm0[3 * k + 1] + R0[3 * i + 2] * Rm0[3 * k + 2];
  const V3<float> pb = p0 - mv(Rb, pm0);
  const V3<float> wb = w0 - mv(Rb, wm0);
  const V3<float> ub = vc0 - cross(w0, mv(R0, load3(L0.c))) - mv(Rb, vm0);

  for (int l0 = 0; l0 < L; l0 += G) {  // every lane takes part in the permutes
    const int l = l0 + g;
    const LinkRec<float> &LL = links_of(Mg)[l < L ? l : L - 1];
    float Rm[9];
    V3<float> pm, wm, vm;
    link_model_state<G>(T, LL, X, v, Rm, pm, wm, vm);
    float Rw[9];
    mm(Rb, Rm, Rw);
    const V3<float> pw = mv(Rb, pm) + pb;
    if (live && l < L) {
      float *out = row + l * io.link_stride;
      store_pose(out, Rw, pw);
      if (vel) {
        const V3<float> ww = wb + mv(Rb, wm);
        const V3<float> vc = ub + cross(wb, pw - p0) + mv(Rb, vm) + cross(ww, mv(Rw, load3(LL.c)));
        out[7] = vc.x; out[8] = vc.y; out[9] = vc.z;
        out[10] = ww.x; out[11] = ww.y; out[12] = ww.z;
      }
    }
  }
}

// Heterogeneous launch: block ranges map to segments (one tree type each).
__global__ __launch_bounds__(256) void k_hetero(HeteroArgs args, int flags) {
  __shared__ float4 s_tab[kTreeTableWords4];
  extern __shared__ float4 s_crba[];  // crba_lds_bytes(largest G) when KF_CRBA, else empty
  int s = 0;
  while (s + 1 < args.nseg && (int64_t)blockIdx.x >= args.seg[s + 1].block0) ++s;
  const HeteroSegment &S = args.seg[s];
  stage_table(s_tab, S.tab);
  __syncthreads();
  const int64_t local = (int64_t)(blockIdx.x - S.block0) * blockDim.x + threadIdx.x;
  const ModelTable<float> *M = reinterpret_cast<const ModelTable<float> *>(s_tab);
  TreeIO io{S.q, S.qd ? S.qd : S.q, S.qdd ? S.qdd : S.q, S.poses, S.J, S.tau, S.M};
  int f = flags;
  if (!S.poses) f &= ~KF_POSES;
  if (!S.J) f &= ~KF_JFULL;
  if (!S.tau) f &= ~KF_RNEA;
  if (!S.M) f &= ~KF_CRBA;
  switch (S.G) {  // uniform per block
    case 8: tree_body<8>(M, S.tab, S.B, local / 8, f, io, s_crba); break;
    case 16: tree_body<16>(M, S.tab, S.B, local / 16, f, io, s_crba); break;
    default: tree_body<32>(M, S.tab, S.B, local / 32, f, io, s_crba); break;
  }
}

}  // namespace

int tree_group(int nj) { return nj < 1 ? 0 : nj <= 8 ? 8 : nj <= 16 ? 16 : nj <= 32 ? 32 : 0; }

hipError_t launch_tree_fused(const ModelTable<float> *tab, int nj, int64_t B, int flags, const float *q,
                             const float *qd, const float *qdd, const KinOut<float> &out, hipStream_t s, int G,
                             hipEvent_t ev0, hipEvent_t ev1) {
  if (B <= 0) return hipSuccess;
  if (!tree_group(nj)) return hipErrorNotSupported;
  if (G == 0) G = tree_group(nj);
  if ((G != 8 && G != 16 && G != 32) || G < nj) return hipErrorInvalidValue;
  const bool use_qd = (flags & KF_RNEA) && !(flags & KF_QD0) && qd, use_qdd = (flags & KF_RNEA) && !(flags & KF_QDD0) && qdd;
  TreeIO io{q, use_qd ? qd : q, use_qdd ? qdd : q, out.poses, out.jfull, out.tau, out.M};
  io.ld_q = out.ld_q;
  io.ld_qd = use_qd ? out.ld_qd : out.ld_q;
  io.ld_qdd = use_qdd ? out.ld_qdd : out.ld_q;
  io.ld_tau = out.ld_tau;
  io.pose_env = out.pose_env;
  io.pose_link = out.pose_link;
  const int64_t per_block = 256 / G;
  dim3 grid((unsigned)((B + per_block - 1) / per_block));
  if (!out.M) flags &= ~KF_CRBA;  // the CRBA tile is allocated only with an M output
  const unsigned lds = (flags & KF_CRBA) ? (unsigned)crba_lds_bytes(G) : 0u;
  if (ev0 || ev1) {  // the kernel's own start and stop recorded into the caller's events
    if (G == 8) hipExtLaunchKernelGGL(k_tree_fused<8>, grid, dim3(256), lds, s, ev0, ev1, 0, tab, B, flags, io);
    else if (G == 16) hipExtLaunchKernelGGL(k_tree_fused<16>, grid, dim3(256), lds, s, ev0, ev1, 0, tab, B, flags, io);
    else hipExtLaunchKernelGGL(k_tree_fused<32>, grid, dim3(256), lds, s, ev0, ev1, 0, tab, B, flags, io);
  } else if (G == 8) {
    hipLaunchKernelGGL(k_tree_fused<8>, grid, dim3(256), lds, s, tab, B, flags, io);
  } else if (G == 16) {
    hipLaunchKernelGGL(k_tree_fused<16>, grid, dim3(256), lds, s, tab, B, flags, io);
  } else {
    hipLaunchKernelGGL(k_tree_fused<32>, grid, dim3(256), lds, s, tab, B, flags, io);
  }
  return hipGetLastError();
}

hipError_t launch_link_state(const ModelTable<float> *tab, int nj, int64_t B, const float *q, int64_t ld_q,
                             const float *qd, int64_t ld_qd, float *body, int64_t env_stride, int64_t link_stride,
                             hipStream_t s) {
  if (B <= 0) return hipSuccess;
  const int G = tree_group(nj);
  if (!G) return hipErrorNotSupported;
  LinkStateIO io{q, qd, ld_q, ld_qd, body, env_stride, link_stride};
  const int64_t per_block = 256 / G;
  dim3 grid((unsigned)((B + per_block - 1) / per_block));
  if (G == 8) hipLaunchKernelGGL(k_link_state<8>, grid, dim3(256), 0, s, tab, B, io);
  else if (G == 16) hipLaunchKernelGGL(k_link_state<16>, grid, dim3(256), 0, s, tab, B, io);
  else hipLaunchKernelGGL(k_link_state<32>, grid, dim3(256), 0, s, tab, B, io);
  return hipGetLastError();
}

hipError_t launch_tree_vjp(const ModelTable<float> *tab, int nj, int64_t B, const float *q, const float *qd,
                           const float *qdd, VjpArgs<float> a, hipStream_t s) {
  if (B <= 0) return hipSuccess;
  const int G = tree_group(nj);
  if (!G) return hipErrorNotSupported;
  const bool dyn = a.gtau || a.tau_star;
  if (!dyn) qd = qdd = q;
  const float *gsrc = a.tau_star ? a.tau_star : a.gtau ? a.gtau : q;
  const int64_t per_block = 256 / G;
  dim3 grid((unsigned)((B + per_block - 1) / per_block));
  if (G == 8) hipLaunchKernelGGL(k_tree_vjp<8>, grid, dim3(256), 0, s, tab, B, q, qd, qdd, gsrc, a);
  else if (G == 16) hipLaunchKernelGGL(k_tree_vjp<16>, grid, dim3(256), 0, s, tab, B, q, qd, qdd, gsrc, a);
  else hipLaunchKernelGGL(k_tree_vjp<32>, grid, dim3(256), 0, s, tab, B, q, qd, qdd, gsrc, a);
  return hipGetLastError();
}

hipError_t launch_hetero(const HeteroArgs &args, int64_t total_blocks, int flags, hipStream_t s) {
  if (total_blocks <= 0) return hipSuccess;
  for (int i = 0; i < args.nseg; ++i) {
    const int G = args.seg[i].G;
    if (!tree_group(args.seg[i].nj) || (G != 8 && G != 16 && G != 32) || G < args.seg[i].nj)
      return hipErrorInvalidValue;
  }
  unsigned lds = 0;
  if (flags & KF_CRBA)
    for (int i = 0; i < args.nseg; ++i)
      if (args.seg[i].M) lds = std::max(lds, (unsigned)crba_lds_bytes(args.seg[i].G));
  hipLaunchKernelGGL(k_hetero, dim3((unsigned)total_blocks), dim3(256), lds, s, args, flags);
  return hipGetLastError();
}

}  // namespace sapamd
