// Lanes-per-joint forward + reverse pass of FK link poses and RNEA for serial
// chains (FP32), the same G lanes x C joints-per-lane layout as the forward
// kernel. One launch recomputes the forward scans and runs the adjoint scans
// (see kernels_adjoint.hip for the formulas):
//   w   prefix sum of g_j col_j                 (virtual velocity)
//   Hw  suffix sum of I_j w_j
//   QS  suffix sum of w_j x* I_j v_j - v_j x* I_j w_j + I_j (v_j x w_j)
//   FP  suffix sum of the link-pose wrenches
// so the gradient costs two more 6-vector scans and one 18-float suffix scan
// on top of the forward pass, all in registers. In loss mode the upstream
// gradients come from loss = sum (tau - tau*)^2 + |p_ee - t|^2 and the
// per-environment loss is reduced across the group into lane 0.
#include <hip/hip_runtime.h>

#include "lanes.h"
#include "launch.h"
#include "spatial.h"

namespace sapamd {

namespace {

using namespace lanes;

constexpr int kMaxChainJoints = 16;
constexpr int kTableWords4 = (int)((kHeaderBytes + sizeof(JointRec<float>) * kMaxChainJoints) / 16);

template <int G, int C>
__global__ __launch_bounds__(256) void k_chain_vjp(const ModelTable<float> *__restrict__ Mg, int64_t B,
                                                   const float *__restrict__ q, const float *__restrict__ qd,
                                                   const float *__restrict__ qdd, const float *__restrict__ gtu_src,
                                                   VjpArgs<float> A) {
  __shared__ float4 s_tab[kTableWords4];
  // 8 lanes x 1 joint: interleaved pairs (joint j of the row's environment p
  // on lane 2j + p), the scans are DPP row shifts with edge defaults.
  constexpr bool IL = G == 8 && C == 1;
  constexpr int lstep = IL ? 2 : 1;  // lane distance between consecutive joint slots
  const int lane = threadIdx.x & 63;
  const int g = IL ? (lane & 15) >> 1 : (int)(threadIdx.x & (G - 1));
  const int lane0 = IL ? ((lane & ~15) | (lane & 1)) : (lane & ~(G - 1));  // the group's joint slot 0
  const int64_t env = IL ? (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) / 16 * 2 + (lane & 1)
                         : (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) / G;
  const bool live = env < B;
  const int64_t e = live ? env : B - 1;
  const int nj = Mg->nj, nv = Mg->nv, L = Mg->nlinks;
  const bool dyn = A.gtau || A.tau_star;
  const bool lossmode = A.tau_star || A.p_target;
  {
    const int n4 = (int)((kHeaderBytes + sizeof(JointRec<float>) * nj) / 16);
    const float4 *src = reinterpret_cast<const float4 *>(Mg);
    for (int i = threadIdx.x; i < n4; i += blockDim.x) s_tab[i] = src[i];
  }
  // ---- dof loads (user order), independent of the table; gtu_src = gtau or tau*
  float qu[C], qdu[C], qddu[C], gu[C];
#pragma unroll
  for (int t = 0; t < C; ++t) {
    const int j = g * C + t;
    const int64_t qi = e * nv + (j < nv ? j : nv - 1);
    qu[t] = q[qi];
    qdu[t] = qd[qi];
    qddu[t] = qdd[qi];
    gu[t] = gtu_src[qi];
  }
  V3<float> ptarget = {0.f, 0.f, 0.f};
  if (A.p_target) ptarget = {A.p_target[e * 3], A.p_target[e * 3 + 1], A.p_target[e * 3 + 2]};
  __syncthreads();
  const ModelTable<float> *M = reinterpret_cast<const ModelTable<float> *>(s_tab);
  const JointRec<float> *J[C];
  bool valid[C];
  float qj[C], qdj[C], qddj[C], gj[C];
#pragma unroll
  for (int t = 0; t < C; ++t) {
    const int j = g * C + t;
    valid[t] = j < nj;
    J[t] = joints_of(M) + (valid[t] ? j : nj - 1);
    const int u = J[t]->vuser;
    if (M->ident_order) {
      qj[t] = qu[t];
      qdj[t] = qdu[t];
      qddj[t] = qddu[t];
      gj[t] = gu[t];
    } else {
      const int src = 4 * (lane0 + lstep * (u / C)), slot = u % C;
      qj[t] = qdj[t] = qddj[t] = gj[t] = 0.f;
#pragma unroll
      for (int s = 0; s < C; ++s) {
        const float a = bperm(src, qu[s]), b = bperm(src, qdu[s]), c = bperm(src, qddu[s]), d = bperm(src, gu[s]);
        qj[t] = slot == s ? a : qj[t];
        qdj[t] = slot == s ? b : qdj[t];
        qddj[t] = slot == s ? c : qddj[t];
        gj[t] = slot == s ? d : gj[t];
      }
    }
    if (!valid[t] || !dyn) qdj[t] = qddj[t] = gj[t] = 0.f;
    if (!valid[t]) qj[t] = 0.f;
  }

  // ---- forward: FK and Jacobian columns
  Xf X[C];
#pragma unroll
  for (int t = 0; t < C; ++t) {
    float s, c;
    joint_sincos(qj[t], &s, &c);
    joint_world(*J[t], qj[t], s, c, static_cast<const float *>(nullptr), V3<float>{0, 0, 0}, true, X[t].R, X[t].p);
    if (!valid[t]) X[t] = identity_xf();
    if (t > 0) X[t] = compose(X[t - 1], X[t]);
  }
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

  float loss = 0.f;
  // ---- link-pose wrenches (one user link per joint: links_le1 is required)
  S6<float> FP[C];
#pragma unroll
  for (int t = 0; t < C; ++t) {
    FP[t] = zero6<float>();
    const int l = J[t]->link;
    if (!valid[t] || l < 0) continue;
    if (A.gposes) {
      float Rw[9], quat[4];
      mm(X[t].R, J[t]->lR, Rw);
      const V3<float> pw = mv(X[t].R, load3(J[t]->lp)) + X[t].p;
      quat_branchless(Rw, quat);
      const float *gp7 = A.gposes + (e * L + l) * 7;
      const V3<float> gp = {gp7[0], gp7[1], gp7[2]}, gq = {gp7[4], gp7[5], gp7[6]};
      const V3<float> u = {quat[1], quat[2], quat[3]};
      const V3<float> tq = 0.5f * (quat[0] * gq - gp7[3] * u + cross(u, gq));
      FP[t] = {gp, cross(pw, gp) + tq};
    } else if (A.p_target && l == A.ee_link) {
      const V3<float> pw = mv(X[t].R, load3(J[t]->lp)) + X[t].p;
      const V3<float> d = pw - ptarget;
      loss += dot(d, d);
      const V3<float> gp = 2.f * d;
      FP[t] = {gp, cross(pw, gp)};
    }
  }

  S6<float> Hw[C], QS[C], v[C], a[C], F[C];
  if (dyn) {
    // ---- forward RNEA
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
    // a[C - 1] takes the inclusive scan, the other slots the exclusive one
    S6<float> Ea = zero6<float>();
    if constexpr (G > 1) {
      S6<float> incl = a[C - 1];
      if constexpr (IL) scan_sum_il(incl);
      else scan_sum<G>(g, incl);
      if constexpr (C > 1) Ea = excl_sum<G>(g, incl);
      a[C - 1] = incl;
    }
    const V3<float> grav = {M->gravity[0], M->gravity[1], M->gravity[2]};
    Ea.l = Ea.l - grav;
    WInertia<float> Y[C];
#pragma unroll
    for (int t = 0; t < C; ++t) {
      if (G > 1 && t == C - 1) a[t].l = a[t].l - grav;
      else a[t] = a[t] + Ea;
      Y[t] = world_inertia_fast(*J[t], X[t].R, X[t].p);
      F[t] = valid[t] ? inertia_apply(Y[t], a[t]) + crossf(v[t], inertia_apply(Y[t], v[t])) : zero6<float>();
    }
#pragma unroll
    for (int t = C - 2; t >= 0; --t) F[t] = F[t] + F[t + 1];
    if constexpr (G > 1) {
      S6<float> incl = F[0];
      if constexpr (IL) scan_suffix_il(incl);
      else scan_suffix<G>(g, incl);
      if constexpr (C > 1) {
        const S6<float> E = excl_suffix<G>(g, incl);
#pragma unroll
        for (int t = 1; t < C; ++t) F[t] = F[t] + E;
      }
      F[0] = incl;
    }
    // ---- upstream gradient on tau, then the virtual velocity w
#pragma unroll
    for (int t = 0; t < C; ++t) {
      const float tj = dot6(col[t], F[t]);
      if (A.tau && live && valid[t]) A.tau[e * nv + J[t]->vuser] = tj;
      if (A.tau_star) {
        const float d = valid[t] ? tj - gj[t] : 0.f;
        loss += d * d;
        gj[t] = 2.f * d;
      }
    }
    S6<float> w[C];
#pragma unroll
    for (int t = 0; t < C; ++t) w[t] = (t > 0 ? w[t - 1] : zero6<float>()) + gj[t] * col[t];
    if constexpr (G > 1) {
      S6<float> incl = w[C - 1];
      if constexpr (IL) scan_sum_il(incl);
      else scan_sum<G>(g, incl);
      if constexpr (C > 1) {
        const S6<float> E = excl_sum<G>(g, incl);
#pragma unroll
        for (int t = 0; t < C - 1; ++t) w[t] = w[t] + E;
      }
      w[C - 1] = incl;
    }
#pragma unroll
    for (int t = 0; t < C; ++t) {
      const S6<float> Iw = inertia_apply(Y[t], w[t]), Iv = inertia_apply(Y[t], v[t]);
      Hw[t] = valid[t] ? Iw : zero6<float>();
      QS[t] = valid[t] ? crossf(w[t], Iv) - crossf(v[t], Iw) + inertia_apply(Y[t], crossm(v[t], w[t])) : zero6<float>();
      // parent-side quantities from the joint's own values
      const float qdk = qdj[t], qddk = qddj[t];
      w[t] = w[t] - gj[t] * col[t];
      a[t] = a[t] - qddk * col[t] - qdk * crossm(v[t], col[t]);
      v[t] = v[t] - qdk * col[t];
    }
    // ---- suffix sums of Hw, QS (and FP below)
#pragma unroll
    for (int t = C - 2; t >= 0; --t) {
      Hw[t] = Hw[t] + Hw[t + 1];
      QS[t] = QS[t] + QS[t + 1];
    }
    if constexpr (G > 1) {
      S6<float> i1 = Hw[0], i2 = QS[0];
      if constexpr (IL) {
        scan_suffix_il(i1);
        scan_suffix_il(i2);
      } else {
        scan_suffix<G>(g, i1);
        scan_suffix<G>(g, i2);
      }
      if constexpr (C > 1) {
        const S6<float> E1 = excl_suffix<G>(g, i1), E2 = excl_suffix<G>(g, i2);
#pragma unroll
        for (int t = 1; t < C; ++t) {
          Hw[t] = Hw[t] + E1;
          QS[t] = QS[t] + E2;
        }
      }
      Hw[0] = i1;
      QS[0] = i2;
    }
    // ---- dqd, dqdd and the RNEA part of dq (w, v, a now hold parent values)
#pragma unroll
    for (int t = 0; t < C; ++t) {
      if (!(live && valid[t])) continue;
      const S6<float> c = col[t];
      const int u = J[t]->vuser;
      if (A.dqdd) A.dqdd[e * nv + u] = dot6(c, Hw[t]);
      if (A.dqd) A.dqd[e * nv + u] = -dot6(c, QS[t] + 2.f * crossf(v[t], Hw[t]));
      FP[t] = FP[t];  // kept separate: summed below
      QS[t] = crossf(v[t], QS[t] + crossf(v[t], Hw[t])) - crossf(w[t], F[t]) - crossf(a[t], Hw[t]);
    }
  }
#pragma unroll
  for (int t = C - 2; t >= 0; --t) FP[t] = FP[t] + FP[t + 1];
  if constexpr (G > 1) {
    S6<float> incl = FP[0];
    if constexpr (IL) scan_suffix_il(incl);
    else scan_suffix<G>(g, incl);
    if constexpr (C > 1) {
      const S6<float> E = excl_suffix<G>(g, incl);
#pragma unroll
      for (int t = 1; t < C; ++t) FP[t] = FP[t] + E;
    }
    FP[0] = incl;
  }
#pragma unroll
  for (int t = 0; t < C; ++t) {
    if (!(live && valid[t])) continue;
    float dq = dot6(col[t], FP[t]);
    if (dyn) dq += dot6(col[t], QS[t]);
    if (A.dq) A.dq[e * nv + J[t]->vuser] = dq;
  }
  if (lossmode && A.loss) {
    if constexpr (IL) scan_suffix1_il(loss);
    else scan_suffix1<G>(g, loss);
    if (live && g == 0) A.loss[e] = loss;
  }
}

template <int G, int C>
void launch_gc(const ModelTable<float> *tab, int64_t B, const float *q, const float *qd, const float *qdd,
               const float *gsrc, VjpArgs<float> a, hipStream_t s) {
  const int64_t envs_per_block = 256 / G;
  dim3 grid((unsigned)((B + envs_per_block - 1) / envs_per_block));
  hipLaunchKernelGGL((k_chain_vjp<G, C>), grid, dim3(256), 0, s, tab, B, q, qd, qdd, gsrc, a);
}

}  // namespace

hipError_t launch_chain_vjp(const ModelTable<float> *tab, int nj, int64_t B, const float *q, const float *qd,
                            const float *qdd, VjpArgs<float> a, hipStream_t s, int G, int C) {
  if (B <= 0) return hipSuccess;
  if (nj < 1 || G * C < nj) return hipErrorNotSupported;
  const bool dyn = a.gtau || a.tau_star;
  if (!dyn) qd = qdd = q;  // read but unused
  const float *gsrc = a.tau_star ? a.tau_star : a.gtau ? a.gtau : q;
#define SAPAMD_SHAPE(g_, c_) \
  if (G == g_ && C == c_) { launch_gc<g_, c_>(tab, B, q, qd, qdd, gsrc, a, s); return hipGetLastError(); }
  SAPAMD_SHAPE(16, 1) SAPAMD_SHAPE(8, 1) SAPAMD_SHAPE(8, 2) SAPAMD_SHAPE(4, 2) SAPAMD_SHAPE(4, 4)
  SAPAMD_SHAPE(2, 4) SAPAMD_SHAPE(2, 8) SAPAMD_SHAPE(1, 8) SAPAMD_SHAPE(1, 16)
#undef SAPAMD_SHAPE
  return hipErrorNotSupported;
}

}  // namespace sapamd
