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
#include <hip/hip_runtime.h>

#include "launch.h"
#include "spatial.h"

namespace sapamd {

namespace {

template <int CTRL> __device__ __forceinline__ float dppf(float x) {
  return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, x), CTRL, 0xF, 0xF, true));
}
// value held by lane g - D of the same 16-lane row (0 for lanes below D)
template <int D> __device__ __forceinline__ float from_below(float x) { return dppf<0x110 + D>(x); }
// value held by lane g + D of the same 16-lane row (0 past the row end)
template <int D> __device__ __forceinline__ float from_above(float x) { return dppf<0x100 + D>(x); }
// value held by wave lane src4 / 4
__device__ __forceinline__ float bperm(int src4, float x) {
  return __builtin_bit_cast(float, __builtin_amdgcn_ds_bpermute(src4, __builtin_bit_cast(int, x)));
}

template <int D> __device__ __forceinline__ S6<float> below6(S6<float> v) {
  return {{from_below<D>(v.l.x), from_below<D>(v.l.y), from_below<D>(v.l.z)},
          {from_below<D>(v.a.x), from_below<D>(v.a.y), from_below<D>(v.a.z)}};
}
template <int D> __device__ __forceinline__ S6<float> above6(S6<float> v) {
  return {{from_above<D>(v.l.x), from_above<D>(v.l.y), from_above<D>(v.l.z)},
          {from_above<D>(v.a.x), from_above<D>(v.a.y), from_above<D>(v.a.z)}};
}
__device__ __forceinline__ S6<float> sel(bool t, S6<float> a, S6<float> b) {
  return {{t ? a.l.x : b.l.x, t ? a.l.y : b.l.y, t ? a.l.z : b.l.z},
          {t ? a.a.x : b.a.x, t ? a.a.y : b.a.y, t ? a.a.z : b.a.z}};
}

struct Xf {
  float R[9];
  V3<float> p;
};
__device__ __forceinline__ Xf compose(const Xf &a, const Xf &b) {
  Xf o;
  mm(a.R, b.R, o.R);
  o.p = mv(a.R, b.p) + a.p;
  return o;
}
__device__ __forceinline__ Xf identity_xf() {
  Xf o;
#pragma unroll
  for (int k = 0; k < 9; ++k) o.R[k] = (k % 4 == 0) ? 1.f : 0.f;
  o.p = {0.f, 0.f, 0.f};
  return o;
}
template <int D> __device__ __forceinline__ Xf below_xf(const Xf &x) {
  Xf o;
#pragma unroll
  for (int k = 0; k < 9; ++k) o.R[k] = from_below<D>(x.R[k]);
  o.p = {from_below<D>(x.p.x), from_below<D>(x.p.y), from_below<D>(x.p.z)};
  return o;
}
__device__ __forceinline__ Xf sel_xf(bool t, const Xf &a, const Xf &b) {
  Xf o;
#pragma unroll
  for (int k = 0; k < 9; ++k) o.R[k] = t ? a.R[k] : b.R[k];
  o.p = {t ? a.p.x : b.p.x, t ? a.p.y : b.p.y, t ? a.p.z : b.p.z};
  return o;
}

// Inclusive prefix product over the group: x_g <- x_0 ... x_g.
template <int G, int D = 1> __device__ __forceinline__ void scan_xf(int g, Xf &x) {
  if constexpr (D < G) {
    Xf y = below_xf<D>(x);
    x = sel_xf(g >= D, compose(y, x), x);
    scan_xf<G, 2 * D>(g, x);
  }
}
template <int G, int D = 1> __device__ __forceinline__ void scan_sum(int g, S6<float> &x) {
  if constexpr (D < G) {
    x = sel(g >= D, x + below6<D>(x), x);
    scan_sum<G, 2 * D>(g, x);
  }
}
template <int G, int D = 1> __device__ __forceinline__ void scan_suffix(int g, S6<float> &x) {
  if constexpr (D < G) {
    x = sel(g + D < G, x + above6<D>(x), x);
    scan_suffix<G, 2 * D>(g, x);
  }
}
// exclusive values from the inclusive scan
template <int G> __device__ __forceinline__ Xf excl_xf(int g, const Xf &incl) {
  if constexpr (G == 1) return identity_xf();
  else return sel_xf(g >= 1, below_xf<1>(incl), identity_xf());
}
template <int G> __device__ __forceinline__ S6<float> excl_sum(int g, S6<float> incl) {
  if constexpr (G == 1) return zero6<float>();
  else return sel(g >= 1, below6<1>(incl), zero6<float>());
}
template <int G> __device__ __forceinline__ S6<float> excl_suffix(int g, S6<float> incl) {
  if constexpr (G == 1) return zero6<float>();
  else return sel(g + 1 < G, above6<1>(incl), zero6<float>());
}

// Shoemake's quaternion with Eigen's branch rule (w first when the trace is
// positive, else the largest diagonal entry), evaluated without divergence.
__device__ __forceinline__ void quat_branchless(const float *R, float *q) {
  float tr = R[0] + R[4] + R[8];
  int i = R[4] > R[0] ? 1 : 0;
  float di = i ? R[4] : R[0];
  int k = tr > 0.f ? 0 : 1 + (R[8] > di ? 2 : i);
  float t = k == 0 ? tr : k == 1 ? R[0] - R[4] - R[8] : k == 2 ? R[4] - R[8] - R[0] : R[8] - R[0] - R[4];
  float r = sqrtf(t + 1.f);
  float h = 0.5f / r;
  float a = (R[7] - R[5]) * h, b = (R[2] - R[6]) * h, c = (R[3] - R[1]) * h;  // w-type
  float d = (R[1] + R[3]) * h, e = (R[2] + R[6]) * h, f = (R[5] + R[7]) * h;  // off-diagonal sums
  float big = 0.5f * r;
  q[0] = k == 0 ? big : k == 1 ? a : k == 2 ? b : c;
  q[1] = k == 0 ? a : k == 1 ? big : k == 2 ? d : e;
  q[2] = k == 0 ? b : k == 1 ? d : k == 2 ? big : f;
  q[3] = k == 0 ? c : k == 1 ? e : k == 2 ? f : big;
}

__device__ __forceinline__ void store_pose(float *P, const float *R, V3<float> p) {
  float qq[4];
  quat_branchless(R, qq);
  P[0] = p.x; P[1] = p.y; P[2] = p.z;
  P[3] = qq[0]; P[4] = qq[1]; P[5] = qq[2]; P[6] = qq[3];
}

// World inertia of a body whose rotational inertia is diagonal in its frame:
// R diag(d) R^T = sum_k d_k r_k r_k^T (columns r_k of R).
__device__ __forceinline__ WInertia<float> world_inertia_fast(const JointRec<float> &J, const float *R, V3<float> p) {
  if (!J.diagI) return world_inertia(J, R, p);
  WInertia<float> Y;
  Y.m = J.m;
  Y.c = mv(R, load3(J.c)) + p;
  const float d0 = J.Ic[0], d1 = J.Ic[3], d2 = J.Ic[5];
  Y.I[0] = d0 * R[0] * R[0] + d1 * R[1] * R[1] + d2 * R[2] * R[2];
  Y.I[1] = d0 * R[0] * R[3] + d1 * R[1] * R[4] + d2 * R[2] * R[5];
  Y.I[2] = d0 * R[0] * R[6] + d1 * R[1] * R[7] + d2 * R[2] * R[8];
  Y.I[3] = d0 * R[3] * R[3] + d1 * R[4] * R[4] + d2 * R[5] * R[5];
  Y.I[4] = d0 * R[3] * R[6] + d1 * R[4] * R[7] + d2 * R[5] * R[8];
  Y.I[5] = d0 * R[6] * R[6] + d1 * R[7] * R[7] + d2 * R[8] * R[8];
  return Y;
}

constexpr int kMaxChainJoints = 16;
constexpr int kTableWords4 = (int)((kHeaderBytes + sizeof(JointRec<float>) * kMaxChainJoints) / 16);

template <int G, int C>
__global__ __launch_bounds__(256) void k_chain_fused(const ModelTable<float> *__restrict__ Mg, int64_t B, int flags,
                                                     const float *__restrict__ q, const float *__restrict__ qd,
                                                     const float *__restrict__ qdd, float *__restrict__ poses,
                                                     float *__restrict__ Jout, float *__restrict__ tau) {
  // model header + joint records staged in LDS once per block
  __shared__ float4 s_tab[kTableWords4];
  const int g = threadIdx.x & (G - 1);
  const int base4 = ((threadIdx.x & 63) & ~(G - 1)) * 4;
  const int64_t env = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) / G;
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
    const int64_t qi = e * nv + (j < nv ? j : nv - 1);
    qu[t] = q[qi];
    qdu[t] = qd[qi];
    qddu[t] = qdd[qi];
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
      const int src = base4 + 4 * (u / C), slot = u % C;
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
    if (!valid[t]) qj[t] = 0.f;
  }

  // ---- FK: local transforms, serial products inside the lane, scan across lanes
  Xf X[C];
#pragma unroll
  for (int t = 0; t < C; ++t) {
    float s, c;
    sincosf(qj[t], &s, &c);
    joint_world(*J[t], qj[t], s, c, static_cast<const float *>(nullptr), V3<float>{0, 0, 0}, true, X[t].R, X[t].p);
    if (!valid[t]) X[t] = identity_xf();
    if (t > 0) X[t] = compose(X[t - 1], X[t]);
  }
  if constexpr (G > 1) {
    Xf incl = X[C - 1];
    scan_xf<G>(g, incl);
    const Xf E = excl_xf<G>(g, incl);
#pragma unroll
    for (int t = 0; t < C; ++t) X[t] = compose(E, X[t]);
  }
  S6<float> col[C];
#pragma unroll
  for (int t = 0; t < C; ++t) col[t] = valid[t] ? jcol(*J[t], X[t].R, X[t].p) : zero6<float>();

  // ---- RNEA
  if (dyn) {
    S6<float> v[C], a[C];
#pragma unroll
    for (int t = 0; t < C; ++t) v[t] = (t > 0 ? v[t - 1] : zero6<float>()) + qdj[t] * col[t];
    if constexpr (G > 1) {
      S6<float> incl = v[C - 1];
      scan_sum<G>(g, incl);
      const S6<float> E = excl_sum<G>(g, incl);
#pragma unroll
      for (int t = 0; t < C; ++t) v[t] = v[t] + E;
    }
#pragma unroll
    for (int t = 0; t < C; ++t)
      a[t] = (t > 0 ? a[t - 1] : zero6<float>()) + qddj[t] * col[t] + qdj[t] * crossm(v[t], col[t]);
    S6<float> Ea = zero6<float>();
    if constexpr (G > 1) {
      S6<float> incl = a[C - 1];
      scan_sum<G>(g, incl);
      Ea = excl_sum<G>(g, incl);
    }
    Ea.l = Ea.l - V3<float>{M->gravity[0], M->gravity[1], M->gravity[2]};
    S6<float> f[C];
#pragma unroll
    for (int t = 0; t < C; ++t) {
      const S6<float> at = a[t] + Ea;
      const WInertia<float> Y = world_inertia_fast(*J[t], X[t].R, X[t].p);
      f[t] = valid[t] ? inertia_apply(Y, at) + crossf(v[t], inertia_apply(Y, v[t])) : zero6<float>();
    }
#pragma unroll
    for (int t = C - 2; t >= 0; --t) f[t] = f[t] + f[t + 1];  // suffix sums inside the lane
    if constexpr (G > 1) {
      S6<float> incl = f[0];
      scan_suffix<G>(g, incl);
      const S6<float> E = excl_suffix<G>(g, incl);
#pragma unroll
      for (int t = 0; t < C; ++t) f[t] = f[t] + E;
    }
#pragma unroll
    for (int t = 0; t < C; ++t)
      if (live && valid[t]) tau[e * nv + J[t]->vuser] = dot6(col[t], f[t]);
  }
  if ((flags & KF_JFULL) && live) {
#pragma unroll
    for (int t = 0; t < C; ++t) {
      if (!valid[t]) continue;
      float *Jb = Jout + e * 6 * nv + J[t]->vuser;
      Jb[0 * nv] = col[t].l.x; Jb[1 * nv] = col[t].l.y; Jb[2 * nv] = col[t].l.z;
      Jb[3 * nv] = col[t].a.x; Jb[4 * nv] = col[t].a.y; Jb[5 * nv] = col[t].a.z;
    }
  }
  // ---- link poses
  if (!(flags & KF_POSES)) return;
  if (M->links_le1) {
    // each joint's lane writes the link attached to it; lane 0 writes the root link
    if (!live) return;
#pragma unroll
    for (int t = 0; t < C; ++t) {
      if (!valid[t] || J[t]->link < 0) continue;
      float Rw[9];
      mm(X[t].R, J[t]->lR, Rw);
      store_pose(poses + (e * L + J[t]->link) * 7, Rw, mv(X[t].R, load3(J[t]->lp)) + X[t].p);
    }
    if (g == 0 && M->root_link >= 0) store_pose(poses + (e * L + M->root_link) * 7, M->root_R, load3(M->root_p));
    return;
  }
  // general link lists: link slot l = l0 + g fetches its joint's pose from the owner lane
  for (int l0 = 0; l0 < L; l0 += G) {
    const int l = l0 + g;
    const LinkRec<float> &LL = links_of(Mg)[l < L ? l : L - 1];  // link records stay in global memory
    const int jl = LL.joint < 0 ? 0 : LL.joint;
    const int src = base4 + 4 * (jl / C), slot = jl % C;
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
    if (live && l < L) store_pose(poses + (e * L + l) * 7, Rw, pw);
  }
}

template <int G, int C>
void launch_gc(const ModelTable<float> *tab, int64_t B, int flags, const float *q, const float *qd, const float *qdd,
               KinOut<float> out, hipStream_t s) {
  const int64_t envs_per_block = 256 / G;
  dim3 grid((unsigned)((B + envs_per_block - 1) / envs_per_block));
  hipLaunchKernelGGL((k_chain_fused<G, C>), grid, dim3(256), 0, s, tab, B, flags, q, qd, qdd, out.poses, out.jfull,
                     out.tau);
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
                              const float *qd, const float *qdd, KinOut<float> out, hipStream_t s, int G, int C) {
  if (B <= 0) return hipSuccess;
  if (nj < 1 || G * C < nj) return hipErrorNotSupported;
  if (!(flags & KF_RNEA) || !qd || !qdd) qd = qdd = q;  // read but unused
#define SAPAMD_SHAPE(g_, c_) \
  if (G == g_ && C == c_) { launch_gc<g_, c_>(tab, B, flags, q, qd, qdd, out, s); return hipGetLastError(); }
  SAPAMD_SHAPE(16, 1) SAPAMD_SHAPE(8, 1) SAPAMD_SHAPE(8, 2) SAPAMD_SHAPE(4, 2) SAPAMD_SHAPE(4, 4)
  SAPAMD_SHAPE(2, 4) SAPAMD_SHAPE(2, 8) SAPAMD_SHAPE(1, 8) SAPAMD_SHAPE(1, 16)
#undef SAPAMD_SHAPE
  return hipErrorNotSupported;
}

}  // namespace sapamd
