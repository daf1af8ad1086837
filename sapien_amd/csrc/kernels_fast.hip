// Lanes-per-joint kernels for serial chains (FP32).
//
// One environment occupies a group of G lanes (G = 1, 2, 4, 8 or 16, inside one
// 16-lane DPP row); lane g owns the C consecutive joints gC .. gC + C - 1
// (G C >= nj). In the world frame every quantity of a serial chain is a scan:
//   oMi[j]  = M_0 M_1 ... M_j               prefix product of SE(3) transforms
//   v[j]    = sum_{k<=j} col_k qd_k          prefix sum
//   a[j]    = -g + sum_{k<=j} (col_k qdd_k + (v_k x col_k) qd_k)
//   F[j]    = sum_{k>=j} f_k                 suffix sum
//   tau[j]  = col_j . F[j]
// Each scan is two-level: serial inside a lane over its C joints, then log2(G)
// Hillis-Steele steps of DPP row shifts across the group (registers only, no
// LDS), then an exclusive fix-up. G = 8, C = 1 spends 8 lanes per environment
// but few serial steps; G = 1, C = nj is the classic one-thread-per-environment
// recursion. The autotuner picks (G, C) per model and batch size.
//
// Memory schedule: every global load is independent and issued at entry (one
// round trip): lane g reads user dofs gC .. gC + C - 1 of q/qd/qdd (an
// environment's dofs are one contiguous run), its C joint records and nothing
// else; the joint -> user-dof permutation is applied in registers with
// ds_bpermute. Each lane then writes the poses of the links attached to its own
// joints (one per joint when the model allows it), its Jacobian columns and
// torques.
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include "lanes.h"
#include "launch.h"
#include "spatial.h"

namespace sapamd {

namespace {

using namespace lanes;

constexpr int kMaxChainJoints = 16;
constexpr int kTableWords4 = (int)((kHeaderBytes + sizeof(JointRec<float>) * kMaxChainJoints) / 16);

// Row strides of the batched arrays, resolved on the host (elements).
// staged: the outputs are contiguous (poses (B, L, 7), J (B, 6, nv), tau
// (B, nv)) and the shape streams them (G >= 8): each wavefront's outputs,
// whose environments are consecutive, go through per-wavefront LDS of
// per_wave floats and leave as contiguous runs, every store instruction
// writing 256 consecutive bytes (whole 64-byte sectors), instead of 28-byte
// pieces of many sectors.
struct Layout {
  int64_t q, qd, qdd, tau, pose_env, pose_link;
  int32_t staged, per_wave;
};

template <int G, int C>
__global__ __launch_bounds__(256) void k_chain_fused(const ModelTable<float> *__restrict__ Mg, int64_t B, int flags,
                                                     const float *__restrict__ q, const float *__restrict__ qd,
                                                     const float *__restrict__ qdd, float *__restrict__ poses,
                                                     float *__restrict__ Jout, float *__restrict__ tau,
                                                     Layout lay) {
  constexpr bool NT = G >= 8;  // streaming output stores (see st())
  // 8 x 1: interleaved pairs (lanes.h), joint j of environment p of a row on
  // lane 2j + p, so the scans are select-free DPP shifts of 2D lanes
  constexpr bool IL = G == 8 && C == 1;
  constexpr int lstep = IL ? 2 : 1;  // lane distance between consecutive joint slots
  // model header + joint records staged in LDS once per block
  __shared__ float4 s_tab[kTableWords4];
  extern __shared__ float s_out[];  // lay.per_wave floats per wavefront when lay.staged
  const int lane = threadIdx.x & 63;
  const int g = IL ? (lane & 15) >> 1 : (int)(threadIdx.x & (G - 1));
  const int lane0 = IL ? ((lane & ~15) | (lane & 1)) : (lane & ~(G - 1));  // the group's joint slot 0
  const int64_t env = IL ? (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) / 16 * 2 + (lane & 1)
                         : (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) / G;
  const bool live = env < B;
  const int64_t e = live ? env : B - 1;
  const int nj = Mg->nj, nv = Mg->nv, L = Mg->nlinks;
  const bool dyn = flags & KF_RNEA;
  {
    const int n4 = (int)((kHeaderBytes + sizeof(JointRec<float>) * nj) / 16);
    const float4 *src = reinterpret_cast<const float4 *>(Mg);
    for (int i = threadIdx.x; i < n4; i += blockDim.x) s_tab[i] = src[i];
  }
  // ---- dof loads, independent of the table (qd/qdd alias q when unused)
  float qu[C], qdu[C], qddu[C];
#pragma unroll
  for (int t = 0; t < C; ++t) {
    const int j = g * C + t;
    const int jc = j < nv ? j : nv - 1;
    qu[t] = q[e * lay.q + jc];
    qdu[t] = qd[e * lay.qd + jc];
    qddu[t] = qdd[e * lay.qdd + jc];
  }
  __syncthreads();
  const ModelTable<float> *M = reinterpret_cast<const ModelTable<float> *>(s_tab);
  const JointRec<float> *J[C];
#pragma unroll
  for (int t = 0; t < C; ++t) {
    const int j = g * C + t;
    J[t] = joints_of(M) + (j < nj ? j : nj - 1);
  }

  // ---- dof values of the lane's joints (user order -> joint order)
  float qj[C], qdj[C], qddj[C];
  bool valid[C];
  const bool ident = M->ident_order;
#pragma unroll
  for (int t = 0; t < C; ++t) {
    valid[t] = g * C + t < nj;
    const int u = J[t]->vuser;
    if (ident) {
      qj[t] = qu[t];
      qdj[t] = qdu[t];
      qddj[t] = qddu[t];
    } else {
      const int src = 4 * (lane0 + lstep * (u / C)), slot = u % C;
      qj[t] = qdj[t] = qddj[t] = 0.f;
#pragma unroll
      for (int s = 0; s < C; ++s) {
        float a = bperm(src, qu[s]), b = bperm(src, qdu[s]), c = bperm(src, qddu[s]);
        qj[t] = slot == s ? a : qj[t];
        qdj[t] = slot == s ? b : qdj[t];
        qddj[t] = slot == s ? c : qddj[t];
      }
    }
    if (!valid[t] || !dyn) qdj[t] = qddj[t] = 0.f;
    if (flags & KF_QD0) qdj[t] = 0.f;
    if (flags & KF_QDD0) qddj[t] = 0.f;
    if (!valid[t]) qj[t] = 0.f;
  }

  // ---- FK: local transforms, serial products inside the lane, scan across lanes
  Xf X[C];
#pragma unroll
  for (int t = 0; t < C; ++t) {
    float s, c;
    joint_sincos(qj[t], &s, &c);
    joint_world(*J[t], qj[t], s, c, static_cast<const float *>(nullptr), V3<float>{0, 0, 0}, true, X[t].R, X[t].p);
    if (!valid[t]) X[t] = identity_xf();
    if (t > 0) X[t] = compose(X[t - 1], X[t]);
  }
  // The inclusive scan is the lane's last joint; only the joints before it
  // need the exclusive prefix (none when C = 1).
  if constexpr (G > 1) {
    Xf incl = X[C - 1];
    if constexpr (IL) scan_xf_il(incl);
    else scan_xf<G>(g, incl);
    if constexpr (C > 1) {
      const Xf E = excl_xf<G>(g, incl);
#pragma unroll
      for (int t = 0; t < C - 1; ++t) X[t] = compose(E, X[t]);
    }
    X[C - 1] = incl;
  }
  S6<float> col[C];
#pragma unroll
  for (int t = 0; t < C; ++t) col[t] = valid[t] ? jcol(*J[t], X[t].R, X[t].p) : zero6<float>();
  // this environment's output rows (64-bit offsets formed once)
  float *const prow = poses + e * (lay.pose_env ? lay.pose_env : (int64_t)L * 7);
  float *const jrow = Jout + e * (int64_t)(6 * nv);
  float *const trow = tau + e * lay.tau;
  const int pl = (int)lay.pose_link;
  // staged outputs: this wavefront's environments e0 .. e0 + nlive - 1, this
  // lane's environment at slot le of the wavefront's LDS runs
  constexpr int EPW = 64 / G;
  const bool staged = NT && lay.staged;
  const int le = IL ? (((lane >> 4) << 1) | (lane & 1)) : lane / G;
  const int64_t e0 = (blockIdx.x * (int64_t)blockDim.x + (threadIdx.x & ~63)) / G;
  const int nlive = (int)(B - e0 < EPW ? (B - e0 > 0 ? B - e0 : 0) : EPW);
  float *const wJ = s_out + (threadIdx.x >> 6) * lay.per_wave;  // [EPW][6][nv]
  float *const wP = wJ + EPW * 6 * nv;                           // [EPW][nj + 1][7]
  float *const wT = wP + EPW * (nj + 1) * 7;                     // [EPW][nv]
  const bool stage_poses = staged && M->links_le1 && L <= nj + 1;

  // ---- Jacobian columns and link poses, emitted before the dynamics so that
  // their stores drain while RNEA computes
  if ((flags & KF_JFULL) && staged) {
#pragma unroll
    for (int t = 0; t < C; ++t) {
      if (!valid[t]) continue;
      float *Jb = wJ + le * 6 * nv + J[t]->vuser;
      Jb[0 * nv] = col[t].l.x; Jb[1 * nv] = col[t].l.y; Jb[2 * nv] = col[t].l.z;
      Jb[3 * nv] = col[t].a.x; Jb[4 * nv] = col[t].a.y; Jb[5 * nv] = col[t].a.z;
    }
  } else if (flags & KF_JFULL) {
#pragma unroll
    for (int t = 0; t < C; ++t) {
      if (!valid[t]) continue;
      if (live) {
        float *Jb = jrow + J[t]->vuser;
        st<NT>(col[t].l.x, Jb + 0 * nv); st<NT>(col[t].l.y, Jb + 1 * nv); st<NT>(col[t].l.z, Jb + 2 * nv);
        st<NT>(col[t].a.x, Jb + 3 * nv); st<NT>(col[t].a.y, Jb + 4 * nv); st<NT>(col[t].a.z, Jb + 5 * nv);
      }
    }
  }
  if ((flags & KF_POSES) && M->links_le1) {
    // each joint's lane writes the link attached to it; lane 0 writes the root link
#pragma unroll
    for (int t = 0; t < C; ++t) {
      if (!valid[t] || J[t]->link < 0) continue;
      float Rw[9];
      V3<float> pw = X[t].p;
      const float *R = X[t].R;
      if (!J[t]->lident) {  // link frame != joint frame
        mm(X[t].R, J[t]->lR, Rw);
        pw = mv(X[t].R, load3(J[t]->lp)) + X[t].p;
        R = Rw;
      }
      if (stage_poses) store_pose<false>(wP + (le * L + J[t]->link) * 7, R, pw);
      else if (live) store_pose<NT>(prow + J[t]->link * pl, R, pw);
    }
    if (g == 0 && M->root_link >= 0) {
      if (stage_poses) store_pose_q<false>(wP + (le * L + M->root_link) * 7, M->root_p, M->root_q);
      else if (live) store_pose_q<NT>(prow + M->root_link * pl, M->root_p, M->root_q);
    }
  } else if (flags & KF_POSES) {
    // general link lists: link slot l = l0 + g fetches its joint's pose from the owner lane
    for (int l0 = 0; l0 < L; l0 += G) {
      const int l = l0 + g;
      const LinkRec<float> &LL = links_of(Mg)[l < L ? l : L - 1];  // link records stay in global memory
      const int jl = LL.joint < 0 ? 0 : LL.joint;
      const int src = 4 * (lane0 + lstep * (jl / C)), slot = jl % C;
      Xf W;
#pragma unroll
      for (int s = 0; s < C; ++s) {
        Xf y;
#pragma unroll
        for (int k = 0; k < 9; ++k) y.R[k] = bperm(src, X[s].R[k]);
        y.p = {bperm(src, X[s].p.x), bperm(src, X[s].p.y), bperm(src, X[s].p.z)};
        W = s == 0 ? y : sel_xf(slot == s, y, W);
      }
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

  if (staged) {  // the wavefront's Jacobian and pose runs, whole sectors per store
    wave_sync();
    if (flags & KF_JFULL) copy_run<NT>(wJ, Jout + e0 * 6 * nv, nlive * 6 * nv, lane, lay.staged & 2);
    if ((flags & KF_POSES) && stage_poses)
      copy_run<NT>(wP, poses + e0 * L * 7, nlive * L * 7, lane, lay.staged & 4);
  }

  // ---- RNEA
  if (!dyn) return;
  S6<float> v[C], a[C];
#pragma unroll
  for (int t = 0; t < C; ++t) v[t] = (t > 0 ? v[t - 1] : zero6<float>()) + qdj[t] * col[t];
  if constexpr (G > 1) {
    S6<float> incl = v[C - 1];
    if constexpr (IL) scan_sum_il(incl);
    else scan_sum<G>(g, incl);
    if constexpr (C > 1) {
      const S6<float> E = excl_sum<G>(g, incl);
#pragma unroll
      for (int t = 0; t < C - 1; ++t) v[t] = v[t] + E;
    }
    v[C - 1] = incl;
  }
#pragma unroll
  for (int t = 0; t < C; ++t)
    a[t] = (t > 0 ? a[t - 1] : zero6<float>()) + qddj[t] * col[t] + qdj[t] * crossm(v[t], col[t]);
  S6<float> Ea = zero6<float>();
  if constexpr (G > 1) {
    S6<float> incl = a[C - 1];
    if constexpr (IL) scan_sum_il(incl);
    else scan_sum<G>(g, incl);
    if constexpr (C > 1) Ea = excl_sum<G>(g, incl);
    a[C - 1] = incl;
  }
  const V3<float> grav = (flags & KF_NOGRAV) ? V3<float>{0.f, 0.f, 0.f}
                                            : V3<float>{M->gravity[0], M->gravity[1], M->gravity[2]};
  S6<float> f[C];
#pragma unroll
  for (int t = 0; t < C; ++t) {
    S6<float> at = (G > 1 && t == C - 1) ? a[t] : a[t] + Ea;
    at.l = at.l - grav;
    const WInertia<float> Y = world_inertia_fast(*J[t], X[t].R, X[t].p);
    f[t] = valid[t] ? inertia_apply(Y, at) + crossf(v[t], inertia_apply(Y, v[t])) : zero6<float>();
  }
#pragma unroll
  for (int t = C - 2; t >= 0; --t) f[t] = f[t] + f[t + 1];  // suffix sums inside the lane
  if constexpr (G > 1) {
    S6<float> incl = f[0];
    if constexpr (IL) scan_suffix_il(incl);
    else scan_suffix<G>(g, incl);
    if constexpr (C > 1) {
      const S6<float> E = excl_suffix<G>(g, incl);
#pragma unroll
      for (int t = 1; t < C; ++t) f[t] = f[t] + E;
    }
    f[0] = incl;
  }
  if (staged) {
#pragma unroll
    for (int t = 0; t < C; ++t)
      if (valid[t]) wT[le * nv + J[t]->vuser] = dot6(col[t], f[t]);
    wave_sync();
    copy_run<NT>(wT, tau + e0 * nv, nlive * nv, lane, lay.staged & 8);
    return;
  }
#pragma unroll
  for (int t = 0; t < C; ++t) {
    if (!valid[t]) continue;
    if (live) st<NT>(dot6(col[t], f[t]), trow + J[t]->vuser);
  }
}

template <int G, int C>
void launch_gc(const ModelTable<float> *tab, int64_t B, int flags, const float *q, const float *qd, const float *qdd,
               KinOut<float> out, hipStream_t s, const Layout &lay, hipEvent_t ev0, hipEvent_t ev1) {
  const int64_t envs_per_block = 256 / G;
  dim3 grid((unsigned)((B + envs_per_block - 1) / envs_per_block));
  const unsigned lds = lay.staged ? 4u * (unsigned)lay.per_wave * (unsigned)sizeof(float) : 0u;  // 4 wavefronts
  if (ev0 || ev1)  // the kernel's own start and stop recorded into the caller's events
    hipExtLaunchKernelGGL((k_chain_fused<G, C>), grid, dim3(256), lds, s, ev0, ev1, 0, tab, B, flags, q, qd, qdd,
                          out.poses, out.jfull, out.tau, lay);
  else
    hipLaunchKernelGGL((k_chain_fused<G, C>), grid, dim3(256), lds, s, tab, B, flags, q, qd, qdd, out.poses,
                       out.jfull, out.tau, lay);
}

}  // namespace

// Supported (G, C) shapes, G lanes per environment x C joints per lane.
static const int kShapes[][2] = {{16, 1}, {8, 1}, {8, 2}, {4, 2}, {4, 4}, {2, 4}, {2, 8}, {1, 8}, {1, 16}};

int chain_shape_count() { return (int)(sizeof(kShapes) / sizeof(kShapes[0])); }
void chain_shape(int i, int *G, int *C) {
  *G = kShapes[i][0];
  *C = kShapes[i][1];
}
bool chain_shape_fits(int G, int C, int nj) { return nj >= 1 && G * C >= nj && (G * C) / 2 < nj + (C > 1 ? C : 0); }

// Default shape before any tuning: 4 lanes x 2 joints up to 8 joints (one
// wave per SIMD at 16k environments), 8 x 2 up to 16.
void default_chain_shape(int nj, int *G, int *C) {
  if (nj <= 8) { *G = 4; *C = 2; }
  else { *G = 8; *C = 2; }
}

int fast_chain_group(int nj) { return nj >= 1 && nj <= 16 ? 1 : 0; }

hipError_t launch_chain_fused(const ModelTable<float> *tab, int nj, int64_t B, int flags, const float *q,
                              const float *qd, const float *qdd, KinOut<float> out, hipStream_t s, int G, int C,
                              hipEvent_t ev0, hipEvent_t ev1) {
  if (B <= 0) return hipSuccess;
  if (nj < 1 || G * C < nj) return hipErrorNotSupported;
  if (!(flags & KF_RNEA) || (flags & KF_QD0) || !qd) qd = q;  // read but unused
  if (!(flags & KF_RNEA) || (flags & KF_QDD0) || !qdd) qdd = q;
  const int nv = nj;  // chains: one dof per joint
  Layout lay;
  lay.q = out.ld_q ? out.ld_q : nv;
  lay.qd = qd == q ? lay.q : (out.ld_qd ? out.ld_qd : nv);
  lay.qdd = qdd == q ? lay.q : (out.ld_qdd ? out.ld_qdd : nv);
  lay.tau = out.ld_tau ? out.ld_tau : nv;
  lay.pose_link = out.pose_link ? out.pose_link : 7;
  lay.pose_env = out.pose_env;  // 0: L * 7, resolved in the kernel from the table
  // staged outputs for the streaming shapes over contiguous outputs (the pose
  // run assumes at most one link per joint plus the root, checked in the kernel)
  lay.staged = G >= 8 && out.pose_env == 0 && lay.pose_link == 7 && lay.tau == nv ? 1 : 0;
  lay.per_wave = (64 / G) * (6 * nv + 7 * (nj + 1) + nv);  // a multiple of 4: the runs stay 16-byte aligned
  // bits 2, 4, 8: the J, pose and tau runs start 16-byte aligned (each
  // wavefront's run starts a multiple of 16 bytes after the array's start)
  auto al16 = [](const void *p) { return ((uintptr_t)p & 15) == 0; };
  if (lay.staged)
    lay.staged |= (al16(out.jfull) ? 2 : 0) | (al16(out.poses) ? 4 : 0) | (al16(out.tau) ? 8 : 0);
#define SAPAMD_SHAPE(g_, c_) \
  if (G == g_ && C == c_) { launch_gc<g_, c_>(tab, B, flags, q, qd, qdd, out, s, lay, ev0, ev1); return hipGetLastError(); }
  SAPAMD_SHAPE(16, 1) SAPAMD_SHAPE(8, 1) SAPAMD_SHAPE(8, 2) SAPAMD_SHAPE(4, 2) SAPAMD_SHAPE(4, 4)
  SAPAMD_SHAPE(2, 4) SAPAMD_SHAPE(2, 8) SAPAMD_SHAPE(1, 8) SAPAMD_SHAPE(1, 16)
#undef SAPAMD_SHAPE
  return hipErrorNotSupported;
}

}  // namespace sapamd
