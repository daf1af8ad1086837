// C ABI implementation (declared in include/sapien_amd.h).
#include <hip/hip_runtime.h>

#include <climits>
#include <cmath>
#include <cstring>
#include <initializer_list>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>
#include <algorithm>
#include <random>

#include "../../include/sapien_amd.h"
#include "launch.h"
#include "model.h"

using namespace sapamd;

namespace {

thread_local std::string g_last_error;

struct DeviceError : std::runtime_error {
  using std::runtime_error::runtime_error;
};

#define HIP_OK(expr)                                                                            \
  do {                                                                                          \
    hipError_t _e = (expr);                                                                     \
    if (_e != hipSuccess) throw DeviceError(std::string(#expr) + ": " + hipGetErrorString(_e)); \
  } while (0)

template <class F> int guarded(F &&f) {
  try {
    f();
    return SAPAMD_OK;
  } catch (const std::invalid_argument &e) {
    g_last_error = e.what();
    return SAPAMD_INVALID_ARGUMENT;
  } catch (const DeviceError &e) {
    g_last_error = e.what();
    return SAPAMD_DEVICE_ERROR;
  } catch (const std::exception &e) {
    g_last_error = e.what();
    return SAPAMD_RUNTIME_ERROR;
  }
}

struct DeviceGuard {
  int prev = -1;
  explicit DeviceGuard(int dev) {
    HIP_OK(hipGetDevice(&prev));
    if (prev != dev) HIP_OK(hipSetDevice(dev));
  }
  ~DeviceGuard() {
    int cur;
    if (hipGetDevice(&cur) == hipSuccess && cur != prev) (void)hipSetDevice(prev);
  }
};

void skew2(const double a[3], double K[9]) {
  K[0] = 0; K[1] = -a[2]; K[2] = a[1];
  K[3] = a[2]; K[4] = 0; K[5] = -a[0];
  K[6] = -a[1]; K[7] = a[0]; K[8] = 0;
}
// Quaternion (w x y z) of a rotation matrix, Shoemake's method with Eigen's
// branch rule (w from the trace when it is positive, else the largest
// diagonal entry first), as Eigen::Quaterniond(R) in pinocchio_model.cpp:159.
void quat_of_matrix(const double *R, double *qq) {
  double t = R[0] + R[4] + R[8];
  if (t > 0) {
    double r = std::sqrt(t + 1), h = 0.5 / r;
    qq[0] = 0.5 * r; qq[1] = (R[7] - R[5]) * h; qq[2] = (R[2] - R[6]) * h; qq[3] = (R[3] - R[1]) * h;
  } else {
    int i = 0;
    if (R[4] > R[0]) i = 1;
    if (R[8] > R[3 * i + i]) i = 2;
    int j = (i + 1) % 3, k2 = (j + 1) % 3;
    double r = std::sqrt(R[3 * i + i] - R[3 * j + j] - R[3 * k2 + k2] + 1), h = 0.5 / r;
    double xyz[3];
    xyz[i] = 0.5 * r;
    xyz[j] = (R[3 * j + i] + R[3 * i + j]) * h;
    xyz[k2] = (R[3 * k2 + i] + R[3 * i + k2]) * h;
    qq[0] = (R[3 * k2 + j] - R[3 * j + k2]) * h;
    qq[1] = xyz[0]; qq[2] = xyz[1]; qq[3] = xyz[2];
  }
}
void mm3(const double *A, const double *B, double *C) {
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) C[3 * i + j] = A[3 * i] * B[j] + A[3 * i + 1] * B[3 + j] + A[3 * i + 2] * B[6 + j];
}

// One stream per device for every model's per-call (one-configuration) work,
// created on first use and kept for the process. Per-call methods synchronize
// before returning, so models can share it. A private segment is reserved per
// hardware queue, for good, from an agent-wide scratch budget shared by all
// queues (hsa_ext_amd.h, HSA_AMD_AGENT_INFO_SCRATCH_LIMIT_MAX/_CURRENT): one
// stream for all models keeps this library's large-private-segment FP64
// kernels on one queue instead of spreading their reservations over every
// queue the streams of many models would map to.
hipStream_t shared_stream(int device) {
  static std::mutex mu;
  static std::vector<hipStream_t> streams;
  std::lock_guard<std::mutex> lock(mu);
  if ((int)streams.size() <= device) streams.resize(device + 1, nullptr);
  if (!streams[device]) {
    DeviceGuard g(device);
    HIP_OK(hipStreamCreate(&streams[device]));
  }
  return streams[device];
}

}  // namespace

struct sapamd_model {
  std::unique_ptr<ArticulatedModel> host;
  int device = 0;
  void *tab32 = nullptr, *tab64 = nullptr;
  int32_t *d_chain = nullptr;   // every link's support chain, concatenated
  std::vector<int> chain_off, chain_len;
  int32_t *d_counter = nullptr;  // CLIK [0] work queue, [1], [3] hand-over counts, [2] second-phase queue
  mutable ClikCont *d_cont = nullptr;  // CLIK hand-over records (grown on demand)
  mutable int64_t cont_cap = 0;
  hipStream_t stream = nullptr;  // the device's shared per-call stream (not owned)
  // per-call buffer: coherent pinned host memory (h_buf) that the kernels read
  // and write directly through its device address (d_buf): no staging copies
  double *h_buf = nullptr, *d_buf = nullptr;
  size_t buf_doubles = 0;
  bool tables_ready = false;
  int variant = SAPAMD_KERNEL_AUTO;
  std::vector<double> pose_cache, jw_cache, jl_cache;
  // Largest cond(M) over sampled configurations (FP64, at refresh) and the
  // precision policy of batched forward dynamics: FP32 ABA, or FP64
  // CRBA + LTDL when the estimate exceeds kCondLimit. FP32 ABA loses about
  // cond(M) * 6e-8 relative (more on floating bases emulated by six 1-dof
  // joints, where the articulated inertias cancel at the root), so above
  // 1e4 it cannot hold a 1e-3 relative tolerance.
  double cond_est = 0;
  int dyn_precision = SAPAMD_DYNAMICS_AUTO;
  static constexpr double kCondLimit = 1e4;
  bool fd64() const {
    return dyn_precision == SAPAMD_DYNAMICS_FP64 || (dyn_precision == SAPAMD_DYNAMICS_AUTO && cond_est > kCondLimit);
  }
  // Rollouts under the lanes-per-joint dynamics: one launch for all steps
  // (default), or one launch per step (the path graph-captured step loops use).
  bool per_step_rollouts = false;
  // Lanes per environment of the lanes-per-joint tree kernels (0 = the
  // smallest of 8, 16, 32 that holds the joints; the autotuner may pick more).
  int tree_G = 0;
  int tree_lanes() const { return tree_G ? tree_G : tree_group(host->njoints()); }
  // Lanes-per-joint forward dynamics: the factorization of M per root subtree
  // (M is block diagonal over them) when that shortens it; rows = its size.
  DynFactorPlan factor_plan() const {
    std::vector<int> parent;
    for (const auto &j : host->joints()) parent.push_back(j.parent);
    return dyn_factor_plan(parent.data(), (int)parent.size());
  }
  int dyn_rows = 0;  // factor_plan().rows of the current tables

  ~sapamd_model() {
    if (stream) (void)hipStreamSynchronize(stream);
    if (tab32) (void)hipFree(tab32);
    if (tab64) (void)hipFree(tab64);
    if (d_chain) (void)hipFree(d_chain);
    if (d_counter) (void)hipFree(d_counter);
    if (d_cont) (void)hipFree(d_cont);
    if (h_buf) (void)hipHostFree(h_buf);
  }

  int nv() const { return host->nv(); }
  int L() const { return host->nlinks(); }

  // The lanes-per-joint CLIK's queue counters and hand-over records for n
  // problems; overlapping CLIK launches on one model race on them.
  void clik_workspace(HeteroClikArgs &A, int64_t n, hipStream_t s) const {
    if (n > cont_cap) {
      DeviceGuard g(device);
      if (d_cont) HIP_OK(hipFree(d_cont));
      d_cont = nullptr;
      cont_cap = 0;
      HIP_OK(hipMalloc(&d_cont, sizeof(ClikCont) * n));
      cont_cap = n;
    }
    A.counter = d_counter;
    A.cont_count = d_counter + 1;
    A.cont_queue = d_counter + 2;
    A.cont_easy = d_counter + 3;
    A.cont = d_cont;
    HIP_OK(launch_fill_i32(d_counter, 0, 4, s));
  }

  template <class T> std::vector<char> build_table() const {
    const auto &J = host->joints();
    const auto &F = host->frames();
    const auto &lf = host->link_frames();
    int nj = (int)J.size(), nl = (int)lf.size();
    std::vector<char> buf(table_bytes<T>(nj, nl), 0);
    auto *hdr = reinterpret_cast<ModelTable<T> *>(buf.data());
    hdr->nj = nj;
    hdr->nlinks = nl;
    hdr->nv = host->nv();
    hdr->chain = host->is_chain() ? 1 : 0;
    hdr->max_depth = host->max_depth();
    hdr->ident_order = 1;
    for (int j = 0; j < nj; ++j)
      if (host->vuser()[j] != j) hdr->ident_order = 0;
    for (int i = 0; i < 3; ++i) hdr->gravity[i] = T(host->gravity()[i]);
    auto *jr = const_cast<JointRec<T> *>(joints_of(hdr));
    std::vector<int> nchild(nj, 0);
    for (int j = 0; j < nj; ++j)
      if (J[j].parent >= 0) nchild[J[j].parent]++;
    for (int j = 0; j < nj; ++j) {
      const Joint &jt = J[j];
      JointRec<T> &r = jr[j];
      double K[9], K2[9], B[9], C[9];
      skew2(jt.axis, K);
      mm3(K, K, K2);
      mm3(jt.place.R, K, B);
      mm3(jt.place.R, K2, C);
      for (int k = 0; k < 9; ++k) {
        r.A[k] = T(jt.place.R[k]);
        r.Bq[k] = jt.kind == JK_PRISMATIC ? T(0) : T(B[k]);
        r.Cq[k] = jt.kind == JK_PRISMATIC ? T(0) : T(C[k]);
      }
      for (int i = 0; i < 3; ++i) {
        r.p[i] = T(jt.place.p[i]);
        r.ap[i] = T(jt.place.R[3 * i] * jt.axis[0] + jt.place.R[3 * i + 1] * jt.axis[1] +
                    jt.place.R[3 * i + 2] * jt.axis[2]);
        r.axis[i] = T(jt.axis[i]);
      }
      double m, c[3], Ic[9];
      jt.inertia.com_form(m, c, Ic);
      r.m = T(m);
      for (int i = 0; i < 3; ++i) r.c[i] = T(c[i]);
      r.Ic[0] = T(Ic[0]); r.Ic[1] = T(Ic[1]); r.Ic[2] = T(Ic[2]);
      r.Ic[3] = T(Ic[4]); r.Ic[4] = T(Ic[5]); r.Ic[5] = T(Ic[8]);
      r.parent = jt.parent;
      r.kind = jt.kind == JK_PRISMATIC ? KIND_PRISMATIC : jt.kind == JK_CONTINUOUS ? KIND_CONTINUOUS : KIND_REVOLUTE;
      r.vuser = host->vuser()[j];
      r.depth = jt.depth;
      r.nchild = nchild[j];
    }
    // per-joint user link (for kernels that emit link poses from the joint's lane)
    std::vector<int> per_joint(nj, 0);
    int on_root = 0;
    hdr->root_link = -1;
    for (int j = 0; j < nj; ++j) {
      jr[j].link = -1;
      jr[j].lident = 0;
      jr[j].diagI = (jr[j].Ic[1] == T(0) && jr[j].Ic[2] == T(0) && jr[j].Ic[4] == T(0)) ? 1 : 0;
    }
    for (int l = 0; l < nl; ++l) {
      const Frame &f = F[lf[l]];
      if (f.joint < 0) {
        ++on_root;
        hdr->root_link = l;
        for (int k = 0; k < 9; ++k) hdr->root_R[k] = T(f.place.R[k]);
        for (int k = 0; k < 3; ++k) hdr->root_p[k] = T(f.place.p[k]);
        double rq[4];
        quat_of_matrix(f.place.R, rq);
        for (int k = 0; k < 4; ++k) hdr->root_q[k] = T(rq[k]);
      } else {
        ++per_joint[f.joint];
        jr[f.joint].link = l;
        bool id = true;
        for (int k = 0; k < 9; ++k) {
          jr[f.joint].lR[k] = T(f.place.R[k]);
          if (f.place.R[k] != (k % 4 == 0 ? 1.0 : 0.0)) id = false;
        }
        for (int k = 0; k < 3; ++k) {
          jr[f.joint].lp[k] = T(f.place.p[k]);
          if (f.place.p[k] != 0.0) id = false;
        }
        jr[f.joint].lident = id ? 1 : 0;
      }
    }
    // pointer-jumping ancestors and depth-first subtree sizes
    for (int j = 0; j < nj; ++j) {
      int a = J[j].parent;
      for (int k = 0; k < 5; ++k) {
        jr[j].anc[k] = a;
        for (int s = 0; s < (1 << k) && a >= 0; ++s) a = J[a].parent;  // advance 2^k more steps
      }
      jr[j].subtree = 1;
    }
    for (int j = nj - 1; j >= 0; --j)
      if (J[j].parent >= 0) jr[J[j].parent].subtree += jr[j].subtree;
    for (int j = 0; j < nj && nj <= 32; ++j)
      for (int a = j; a >= 0; a = J[a].parent) jr[j].ancmask |= 1u << a;
    hdr->links_le1 = on_root <= 1;
    for (int j = 0; j < nj; ++j)
      if (per_joint[j] > 1) hdr->links_le1 = 0;
    auto *lr = const_cast<LinkRec<T> *>(links_of(hdr));
    for (int l = 0; l < nl; ++l) {
      const Frame &f = F[lf[l]];
      bool ident = true;
      for (int k = 0; k < 9; ++k) {
        lr[l].R[k] = T(f.place.R[k]);
        if (f.place.R[k] != (k % 4 == 0 ? 1.0 : 0.0)) ident = false;
      }
      for (int i = 0; i < 3; ++i) {
        lr[l].p[i] = T(f.place.p[i]);
        if (f.place.p[i] != 0.0) ident = false;
      }
      for (int i = 0; i < 3; ++i) lr[l].c[i] = T(f.com[i]);
      lr[l].joint = f.joint;
      lr[l].ident = ident ? 1 : 0;
    }
    return buf;
  }

  // Rebuild device tables and per-call buffers after the orders change.
  void refresh() {
    tables_ready = false;
    if (!host->supported() || !host->joint_order_set()) return;
    cond_est = estimate_condition();
    if (device < 0) return;
    if (host->njoints() > kMaxGenericJoints)
      throw std::runtime_error("models with more than 64 movable joints are not supported");
    DeviceGuard g(device);
    HIP_OK(hipStreamSynchronize(stream));
    auto t32 = build_table<float>();
    auto t64 = build_table<double>();
    dyn_rows = factor_plan().rows;
    if (tab32) HIP_OK(hipFree(tab32));
    if (tab64) HIP_OK(hipFree(tab64));
    tab32 = tab64 = nullptr;
    HIP_OK(hipMalloc(&tab32, t32.size()));
    HIP_OK(hipMalloc(&tab64, t64.size()));
    HIP_OK(hipMemcpy(tab32, t32.data(), t32.size(), hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(tab64, t64.data(), t64.size(), hipMemcpyHostToDevice));
    std::vector<int32_t> chains;
    chain_off.assign(L(), 0);
    chain_len.assign(L(), 0);
    for (int l = 0; l < L(); ++l) {
      int jl = host->frames()[host->link_frames()[l]].joint;
      std::vector<int> c = jl >= 0 ? host->support(jl) : std::vector<int>{};
      chain_off[l] = (int)chains.size();
      chain_len[l] = (int)c.size();
      chains.insert(chains.end(), c.begin(), c.end());
    }
    if (d_chain) HIP_OK(hipFree(d_chain));
    d_chain = nullptr;
    HIP_OK(hipMalloc(&d_chain, sizeof(int32_t) * (chains.size() + 1)));
    if (!chains.empty())
      HIP_OK(hipMemcpy(d_chain, chains.data(), chains.size() * sizeof(int32_t), hipMemcpyHostToDevice));
    size_t n = nv(), l = L();
    size_t need = 8 * n + 7 * l + 12 * n * l + 2 * n * n + 64;
    if (need > buf_doubles) {
      if (h_buf) HIP_OK(hipHostFree(h_buf));
      d_buf = nullptr;
      h_buf = nullptr;
      HIP_OK(hipHostMalloc(&h_buf, need * sizeof(double), hipHostMallocCoherent));
      HIP_OK(hipHostGetDevicePointer(reinterpret_cast<void **>(&d_buf), h_buf, 0));
      buf_doubles = need;
    }
    tables_ready = true;
  }

  // cond(M) = lambda_max / lambda_min over 64 (16 above 32 dofs) seeded configurations drawn
  // uniformly within the joint limits (continuous joints in [-pi, pi]). Host
  // FP64 composite-rigid-body mass matrices (no GPU work at set-up), then
  // eigenvalues by cyclic Jacobi rotations.
  double estimate_condition() const {
    const int n = nv(), S = n > 32 ? 16 : 64;
    if (n == 0) return 1.0;
    const auto &J = host->joints();
    const int nj = (int)J.size();
    std::vector<double> Mh((size_t)S * n * n, 0.0);
    std::mt19937_64 rng(0x5eed);
    std::uniform_real_distribution<double> U(0.0, 1.0);
    std::vector<double> R(9 * nj), P(3 * nj), col(6 * nj), Ym(nj), Yh(3 * nj), YI(9 * nj);
    for (int b = 0; b < S; ++b) {
      for (int j = 0; j < nj; ++j) {
        const Joint &jt = J[j];
        const bool cont = jt.kind == JK_CONTINUOUS;
        const double lo = cont ? -M_PI : jt.lower, hi = cont ? M_PI : jt.upper;
        const double q = lo + (hi - lo) * U(rng);
        // local transform: place * (rotation about / translation along axis)
        double Rl[9], pl[3], K[9], K2[9];
        skew2(jt.axis, K);
        mm3(K, K, K2);
        const double sq = jt.kind == JK_PRISMATIC ? 0.0 : std::sin(q), cq = jt.kind == JK_PRISMATIC ? 1.0 : std::cos(q);
        double Rq[9];
        for (int k = 0; k < 9; ++k) Rq[k] = (k % 4 == 0 ? 1.0 : 0.0) + sq * K[k] + (1 - cq) * K2[k];
        mm3(jt.place.R, Rq, Rl);
        for (int i = 0; i < 3; ++i) {
          pl[i] = jt.place.p[i];
          if (jt.kind == JK_PRISMATIC)
            for (int k = 0; k < 3; ++k) pl[i] += q * jt.place.R[3 * i + k] * jt.axis[k];
        }
        double *Rj = &R[9 * j], *pj = &P[3 * j];
        if (jt.parent < 0) {
          std::copy(Rl, Rl + 9, Rj);
          std::copy(pl, pl + 3, pj);
        } else {
          const double *Rp = &R[9 * jt.parent], *pp = &P[3 * jt.parent];
          mm3(Rp, Rl, Rj);
          for (int i = 0; i < 3; ++i) pj[i] = pp[i] + Rp[3 * i] * pl[0] + Rp[3 * i + 1] * pl[1] + Rp[3 * i + 2] * pl[2];
        }
        double a[3];
        for (int i = 0; i < 3; ++i) a[i] = Rj[3 * i] * jt.axis[0] + Rj[3 * i + 1] * jt.axis[1] + Rj[3 * i + 2] * jt.axis[2];
        double *c6 = &col[6 * j];
        if (jt.kind == JK_PRISMATIC) {
          c6[0] = a[0]; c6[1] = a[1]; c6[2] = a[2]; c6[3] = c6[4] = c6[5] = 0;
        } else {
          c6[0] = pj[1] * a[2] - pj[2] * a[1];
          c6[1] = pj[2] * a[0] - pj[0] * a[2];
          c6[2] = pj[0] * a[1] - pj[1] * a[0];
          c6[3] = a[0]; c6[4] = a[1]; c6[5] = a[2];
        }
        // inertia about the world origin: m, h = m c_w, I_O = R Ic R^T + m(|c|^2 1 - c c^T)
        double m, c[3], Ic[9], cw[3], T1[9], Iw[9];
        jt.inertia.com_form(m, c, Ic);
        for (int i = 0; i < 3; ++i) cw[i] = pj[i] + Rj[3 * i] * c[0] + Rj[3 * i + 1] * c[1] + Rj[3 * i + 2] * c[2];
        mm3(Rj, Ic, T1);
        for (int i = 0; i < 3; ++i)
          for (int k = 0; k < 3; ++k) Iw[3 * i + k] = T1[3 * i] * Rj[3 * k] + T1[3 * i + 1] * Rj[3 * k + 1] + T1[3 * i + 2] * Rj[3 * k + 2];
        const double cc = cw[0] * cw[0] + cw[1] * cw[1] + cw[2] * cw[2];
        Ym[j] = m;
        for (int i = 0; i < 3; ++i) {
          Yh[3 * j + i] = m * cw[i];
          for (int k = 0; k < 3; ++k) YI[9 * j + 3 * i + k] = Iw[3 * i + k] + m * ((i == k ? cc : 0.0) - cw[i] * cw[k]);
        }
      }
      for (int j = nj - 1; j >= 0; --j) {  // composite inertias (children follow parents)
        const int p = J[j].parent;
        if (p < 0) continue;
        Ym[p] += Ym[j];
        for (int i = 0; i < 3; ++i) Yh[3 * p + i] += Yh[3 * j + i];
        for (int i = 0; i < 9; ++i) YI[9 * p + i] += YI[9 * j + i];
      }
      double *Mb = &Mh[(size_t)b * n * n];
      for (int j = 0; j < nj; ++j) {
        const double *l = &col[6 * j], *w = l + 3, *h = &Yh[3 * j], *I = &YI[9 * j];
        // f = Yc_j col_j: f_l = m l + w x h, f_a = h x l + I_O w
        double f[6];
        f[0] = Ym[j] * l[0] + w[1] * h[2] - w[2] * h[1];
        f[1] = Ym[j] * l[1] + w[2] * h[0] - w[0] * h[2];
        f[2] = Ym[j] * l[2] + w[0] * h[1] - w[1] * h[0];
        for (int i = 0; i < 3; ++i) f[3 + i] = I[3 * i] * w[0] + I[3 * i + 1] * w[1] + I[3 * i + 2] * w[2];
        f[3] += h[1] * l[2] - h[2] * l[1];
        f[4] += h[2] * l[0] - h[0] * l[2];
        f[5] += h[0] * l[1] - h[1] * l[0];
        const int uj = host->vuser()[j];
        for (int i = j; i >= 0; i = J[i].parent) {
          const double *ci = &col[6 * i];
          double v = 0;
          for (int k = 0; k < 6; ++k) v += ci[k] * f[k];
          const int ui = host->vuser()[i];
          Mb[ui * n + uj] = v;
          Mb[uj * n + ui] = v;
        }
      }
    }
    double worst = 1.0;
    std::vector<double> A((size_t)n * n);
    for (int b = 0; b < S; ++b) {
      std::copy(Mh.begin() + (size_t)b * n * n, Mh.begin() + (size_t)(b + 1) * n * n, A.begin());
      for (int sweep = 0; sweep < 50; ++sweep) {
        double off = 0;
        for (int i = 0; i < n; ++i)
          for (int k = i + 1; k < n; ++k) off += A[i * n + k] * A[i * n + k];
        if (off < 1e-30) break;
        for (int i = 0; i < n; ++i)
          for (int k = i + 1; k < n; ++k) {
            const double aik = A[i * n + k];
            if (std::abs(aik) < 1e-300) continue;
            const double th = (A[k * n + k] - A[i * n + i]) / (2 * aik);
            const double t = (th >= 0 ? 1.0 : -1.0) / (std::abs(th) + std::sqrt(th * th + 1));
            const double c = 1 / std::sqrt(t * t + 1), sn = t * c;
            for (int r = 0; r < n; ++r) {  // columns i, k
              const double ari = A[r * n + i], ark = A[r * n + k];
              A[r * n + i] = c * ari - sn * ark;
              A[r * n + k] = sn * ari + c * ark;
            }
            for (int r = 0; r < n; ++r) {  // rows i, k
              const double air = A[i * n + r], akr = A[k * n + r];
              A[i * n + r] = c * air - sn * akr;
              A[k * n + r] = sn * air + c * akr;
            }
          }
      }
      double lo = A[0], hi = A[0];
      for (int i = 1; i < n; ++i) {
        lo = std::min(lo, A[i * n + i]);
        hi = std::max(hi, A[i * n + i]);
      }
      worst = std::max(worst, lo > 0 ? hi / lo : INFINITY);
    }
    return worst;
  }

  void reset_caches() {
    // pinocchio::Data starts with oMi = identity and J = 0: a link pose read
    // before any compute is the link's placement in its joint frame.
    size_t n = nv(), l = L();
    pose_cache.assign(7 * l, 0.0);
    for (size_t k = 0; k < l; ++k) {
      const Frame &f = host->frames()[host->link_frames()[k]];
      const double *R = f.place.R;
      for (int i = 0; i < 3; ++i) pose_cache[7 * k + i] = f.place.p[i];
      quat_of_matrix(R, &pose_cache[7 * k + 3]);
    }
    jw_cache.assign(6 * n * l, 0.0);
    jl_cache.assign(6 * n * l, 0.0);
  }

  void require_tables() const {
    if (device < 0) throw DeviceError("no HIP device is available");
    if (!host->supported())
      throw std::runtime_error("Unsupported joint in computation. Currently support: fixed, revolute, prismatic");
    if (!tables_ready) throw std::runtime_error("posS2P failed");
  }
  void check_link(int32_t link) const {
    if (link < 0 || link >= L()) throw std::runtime_error("link index out of bound");
  }
  void check_vec(int64_t len, const char *what) const {
    if (len != nv()) throw std::runtime_error(std::string(what) + " must have length nv");
  }

  // arrays of n doubles into the per-call buffer, back to back from `off` (in
  // doubles); returns their device address. Every per-call method ends in
  // fetch(), so no kernel still reads the buffer when the next call writes it.
  double *stage_seq(std::initializer_list<const double *> srcs, size_t n, size_t off) {
    size_t k = 0;
    for (const double *src : srcs) std::memcpy(h_buf + off + n * k++, src, n * sizeof(double));
    return d_buf + off;
  }
  double *stage(const double *src, size_t n, size_t off) {
    std::memcpy(h_buf + off, src, n * sizeof(double));
    return d_buf + off;
  }
  // host view of a device address inside the per-call buffer
  template <class T> T *host_of(T *dev) const {
    return reinterpret_cast<T *>(reinterpret_cast<char *>(h_buf) +
                                 (reinterpret_cast<char *>(dev) - reinterpret_cast<char *>(d_buf)));
  }
  // wait for the per-call kernel, then copy n doubles at `off` to dst (dst =
  // nullptr: the caller reads h_buf)
  void fetch(double *dst, size_t n, size_t off) {
    HIP_OK(hipStreamSynchronize(stream));
    if (dst) std::memcpy(dst, h_buf + off, n * sizeof(double));
  }
  // serial chain of <= 16 joints: the G x C chain kernels apply
  bool chain_kernel() const { return host->njoints() >= 1 && host->is_chain() && host->njoints() <= 16; }
  int effective_variant() const {
    bool fast_ok = chain_kernel() || tree_group(host->njoints()) > 0;
    if (variant == SAPAMD_KERNEL_THREAD_PER_ENV) return SAPAMD_KERNEL_THREAD_PER_ENV;
    if (variant == SAPAMD_KERNEL_LANES_PER_JOINT) {
      if (!fast_ok) throw std::runtime_error("lanes-per-joint kernels need a tree of at most 32 joints");
      return SAPAMD_KERNEL_LANES_PER_JOINT;
    }
    return fast_ok ? SAPAMD_KERNEL_LANES_PER_JOINT : SAPAMD_KERNEL_THREAD_PER_ENV;
  }
  // Chain kernel shape (G lanes per environment x C joints per lane) per batch
  // size bucket: tuned entries override the default.
  struct ShapeEntry {
    int64_t max_batch;
    int G, C;
  };
  std::vector<ShapeEntry> shapes;
  void shape_for(int64_t B, int *G, int *C) const {
    for (const auto &s : shapes)
      if (B <= s.max_batch) {
        *G = s.G;
        *C = s.C;
        return;
      }
    default_chain_shape(host->njoints(), G, C);
  }
  const ModelTable<double> *t64() const { return static_cast<const ModelTable<double> *>(tab64); }
  const ModelTable<float> *t32() const { return static_cast<const ModelTable<float> *>(tab32); }
};

extern "C" {

const char *sapamd_last_error(void) { return g_last_error.c_str(); }
const char *sapamd_version(void) { return "sapien_amd 0.1.0 (gfx950)"; }

int sapamd_model_create(const char *urdf, const double gravity[3], sapamd_model **out) {
  return guarded([&] {
    if (!urdf || !gravity || !out) throw std::invalid_argument("null argument");
    auto m = std::make_unique<sapamd_model>();
    m->host = std::make_unique<ArticulatedModel>(std::string(urdf), gravity);
    // Without a HIP device the model still parses and orders (host logic);
    // every compute call then fails with a device error.
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
      (void)hipGetLastError();
      m->device = -1;
    } else {
      HIP_OK(hipGetDevice(&m->device));
      m->stream = shared_stream(m->device);
      HIP_OK(hipMalloc(&m->d_counter, sizeof(int32_t) * 4));
    }
    m->reset_caches();
    *out = m.release();
  });
}

void sapamd_model_destroy(sapamd_model *m) { delete m; }

int sapamd_model_info(const sapamd_model *m, sapamd_info *out) {
  return guarded([&] {
    if (!m || !out) throw std::invalid_argument("null argument");
    out->nv = m->host->nv();
    out->nq = m->host->nq();
    out->njoints = m->host->njoints();
    out->nlinks = m->host->nlinks();
    out->max_depth = m->host->max_depth();
    out->chain = m->host->is_chain();
    out->supported = m->host->supported();
    out->order_set = m->host->joint_order_set();
    out->device = m->device;
  });
}

int sapamd_set_joint_order(sapamd_model *m, const char *const *names, int32_t n) {
  return guarded([&] {
    std::vector<std::string> v(names, names + n);
    m->host->set_joint_order(v);
    m->refresh();
    m->reset_caches();
  });
}

int sapamd_set_link_order(sapamd_model *m, const char *const *names, int32_t n) {
  return guarded([&] {
    std::vector<std::string> v(names, names + n);
    m->host->set_link_order(v);
    m->refresh();
    m->reset_caches();
  });
}

int sapamd_joint_limits(const sapamd_model *m, double *lower, double *upper) {
  return guarded([&] {
    if (!m->host->joint_order_set()) throw std::runtime_error("posS2P failed");
    const auto &J = m->host->joints();
    for (size_t j = 0; j < J.size(); ++j) {
      int u = m->host->vuser()[j];
      if (u < 0) continue;
      bool cont = J[j].kind == JK_CONTINUOUS;
      lower[u] = cont ? -M_PI : J[j].lower;
      upper[u] = cont ? M_PI : J[j].upper;
    }
  });
}

int sapamd_continuous_mask(const sapamd_model *m, int32_t *out) {
  return guarded([&] {
    if (!m->host->joint_order_set()) throw std::runtime_error("posS2P failed");
    const auto &J = m->host->joints();
    for (size_t j = 0; j < J.size(); ++j) {
      int u = m->host->vuser()[j];
      if (u >= 0) out[u] = J[j].kind == JK_CONTINUOUS;
    }
  });
}

// ---------------------------------------------------------------- per-call
int sapamd_compute_forward_kinematics(sapamd_model *m, const double *qpos, int64_t n) {
  return guarded([&] {
    m->host->check_compute((size_t)n);
    m->require_tables();
    DeviceGuard g(m->device);
    size_t nv = m->nv(), L = m->L();
    double *dq = m->stage(qpos, nv, 0);
    KinOut<double> o;
    o.poses = m->d_buf + nv;
    HIP_OK(launch_kinematics<double>(m->t64(), m->host->njoints(), 1, KF_POSES, dq, nullptr, nullptr, o, m->stream));
    m->fetch(m->pose_cache.data(), 7 * L, nv);
  });
}

int sapamd_get_link_pose(const sapamd_model *m, int32_t link, float pose7[7]) {
  return guarded([&] {
    m->check_link(link);
    for (int i = 0; i < 7; ++i) pose7[i] = (float)m->pose_cache[7 * link + i];
  });
}

int sapamd_compute_full_jacobian(sapamd_model *m, const double *qpos, int64_t n) {
  return guarded([&] {
    m->host->check_compute((size_t)n);
    m->require_tables();
    DeviceGuard g(m->device);
    size_t nv = m->nv(), L = m->L();
    double *dq = m->stage(qpos, nv, 0);
    KinOut<double> o;
    o.poses = m->d_buf + nv;
    o.jw_links = o.poses + 7 * L;
    o.jl_links = o.jw_links + 6 * nv * L;
    HIP_OK(launch_kinematics<double>(m->t64(), m->host->njoints(), 1, KF_POSES | KF_JLINKS, dq, nullptr, nullptr, o,
                                     m->stream));
    // poses and both link Jacobian sets are contiguous in the device buffer:
    // one copy and one wait for the three caches
    m->fetch(nullptr, 7 * L + 12 * nv * L, nv);
    std::memcpy(m->pose_cache.data(), m->h_buf + nv, 7 * L * sizeof(double));
    std::memcpy(m->jw_cache.data(), m->h_buf + nv + 7 * L, 6 * nv * L * sizeof(double));
    std::memcpy(m->jl_cache.data(), m->h_buf + nv + 7 * L + 6 * nv * L, 6 * nv * L * sizeof(double));
  });
}

int sapamd_get_link_jacobian(const sapamd_model *m, int32_t link, int32_t local, double *out) {
  return guarded([&] {
    m->check_link(link);
    size_t sz = 6 * (size_t)m->nv();
    const auto &c = local ? m->jl_cache : m->jw_cache;
    std::memcpy(out, c.data() + link * sz, sz * sizeof(double));
  });
}

int sapamd_compute_single_link_local_jacobian(sapamd_model *m, const double *qpos, int64_t n, int32_t link,
                                              double *out) {
  return guarded([&] {
    m->check_link(link);
    m->host->check_compute((size_t)n);
    m->require_tables();
    DeviceGuard g(m->device);
    size_t nv = m->nv();
    double *dq = m->stage(qpos, nv, 0);
    HIP_OK(launch_link_local_jacobian<double>(m->t64(), m->host->njoints(), 1, link, dq, m->d_buf + nv, m->stream));
    m->fetch(out, 6 * nv, nv);
  });
}

int sapamd_compute_generalized_mass_matrix(sapamd_model *m, const double *qpos, int64_t n, double *out) {
  return guarded([&] {
    m->host->check_compute((size_t)n);
    m->require_tables();
    DeviceGuard g(m->device);
    size_t nv = m->nv();
    double *dq = m->stage(qpos, nv, 0);
    HIP_OK(launch_crba<double>(m->t64(), m->host->njoints(), 1, dq, m->d_buf + nv, m->stream));
    m->fetch(out, nv * nv, nv);
  });
}

int sapamd_compute_coriolis_matrix(sapamd_model *m, const double *qpos, const double *qvel, int64_t n,
                                   double *out) {
  return guarded([&] {
    m->host->check_compute((size_t)n);
    m->require_tables();
    DeviceGuard g(m->device);
    size_t nv = m->nv();
    double *dq = m->stage_seq({qpos, qvel}, nv, 0), *dv = dq + nv;
    HIP_OK(launch_coriolis<double>(m->t64(), m->host->njoints(), 1, dq, dv, m->d_buf + 2 * nv, m->stream));
    m->fetch(out, nv * nv, 2 * nv);
  });
}

int sapamd_compute_inverse_dynamics(sapamd_model *m, const double *qpos, const double *qvel, const double *qacc,
                                    int64_t n, double *out) {
  return guarded([&] {
    m->host->check_compute((size_t)n);
    m->require_tables();
    DeviceGuard g(m->device);
    size_t nv = m->nv();
    double *dq = m->stage_seq({qpos, qvel, qacc}, nv, 0), *dv = dq + nv, *da = dq + 2 * nv;
    KinOut<double> o;
    o.tau = m->d_buf + 3 * nv;
    HIP_OK(launch_kinematics<double>(m->t64(), m->host->njoints(), 1, KF_RNEA, dq, dv, da, o, m->stream));
    m->fetch(out, nv, 3 * nv);
  });
}

int sapamd_compute_forward_dynamics(sapamd_model *m, const double *qpos, const double *qvel, const double *qf,
                                    int64_t n, double *out) {
  return guarded([&] {
    m->host->check_compute((size_t)n);
    m->require_tables();
    DeviceGuard g(m->device);
    size_t nv = m->nv();
    double *dq = m->stage_seq({qpos, qvel, qf}, nv, 0), *dv = dq + nv, *df = dq + 2 * nv;
    HIP_OK(launch_aba<double>(m->t64(), m->host->njoints(), 1, dq, dv, df, m->d_buf + 3 * nv, m->stream));
    m->fetch(out, nv, 3 * nv);
  });
}

int sapamd_compute_inverse_kinematics(sapamd_model *m, int32_t link, const float pose7[7],
                                      const double *initial_qpos, int64_t n0, const int32_t *active_qmask,
                                      int64_t nm, double eps, int32_t max_iterations, double dt, double damp,
                                      double *q_out, int32_t *success, double err6[6]) {
  return guarded([&] {
    m->check_link(link);
    if (n0 > 0) m->host->check_compute((size_t)n0);
    if (!m->host->supported())
      throw std::runtime_error("Unsupported joint in computation. Currently support: fixed, revolute, prismatic");
    if (!m->host->joint_order_set()) throw std::runtime_error("posP2S failed");
    if (nm > 0 && nm != m->nv()) throw std::runtime_error("active_qmask must have length nv");
    m->require_tables();
    DeviceGuard g(m->device);
    size_t nv = m->nv();
    // device buffer: [target 7 | q0 nv | q_out nv | err 6 | mask nv (int) | success, iters (int)]
    double tgt[7];
    for (int i = 0; i < 7; ++i) tgt[i] = pose7[i];
    double *dt_ = m->stage(tgt, 7, 0);
    double *dq0 = n0 > 0 ? m->stage(initial_qpos, nv, 8) : nullptr;
    double *dqo = m->d_buf + 8 + nv;
    double *derr = dqo + nv;
    int32_t *dmask = reinterpret_cast<int32_t *>(derr + 8);
    int32_t *dints = dmask + nv + 2;
    if (nm > 0) std::memcpy(m->host_of(dmask), active_qmask, nv * sizeof(int32_t));
    ClikParams P;
    P.link = link;
    P.chain = m->d_chain + m->chain_off[link];
    P.nchain = m->chain_len[link];
    P.eps = eps;
    P.dt = dt;
    P.damp = damp;
    P.max_iter = max_iterations;
    P.mask = nm > 0 ? dmask : nullptr;
    P.mask_ld = 0;
    P.counter = m->d_counter;
    P.persistent_waves = 0;
    HIP_OK(launch_clik<double>(m->t64(), m->host->njoints(), 1, P, dt_, dq0, dqo, dints, derr, dints + 1,
                               m->stream));
    m->fetch(q_out, nv, 8 + nv);
    m->fetch(err6, 6, 8 + 2 * nv);
    *success = *m->host_of(dints);
  });
}

// ---------------------------------------------------------------- batched
namespace {
void check_batched(const sapamd_model *m, int64_t B) {
  if (!m) throw std::invalid_argument("null model");
  if (B < 0) throw std::invalid_argument("negative batch size");
  m->require_tables();
}
void need(const void *p, const char *what) {
  if (!p) throw std::invalid_argument(std::string(what) + " must not be null");
}
hipStream_t S(void *s) { return static_cast<hipStream_t>(s); }
}  // namespace

int sapamd_set_kernel_variant(sapamd_model *m, int32_t variant) {
  return guarded([&] {
    if (variant < 0 || variant > 2) throw std::invalid_argument("unknown kernel variant");
    int old = m->variant;
    m->variant = variant;
    try {
      (void)m->effective_variant();
    } catch (...) {
      m->variant = old;
      throw;
    }
  });
}

int sapamd_get_kernel_variant(const sapamd_model *m, int32_t *variant) {
  return guarded([&] { *variant = m->effective_variant(); });
}

int sapamd_chain_shapes(int32_t *G, int32_t *C, int32_t cap) {
  int n = chain_shape_count();
  for (int i = 0; i < n && i < cap; ++i) {
    int g, c;
    chain_shape(i, &g, &c);
    G[i] = g;
    C[i] = c;
  }
  return n;
}

int sapamd_set_chain_shape(sapamd_model *m, int64_t max_batch, int32_t G, int32_t C) {
  return guarded([&] {
    if (G == 0 && C == 0) {  // clear all tuned entries
      m->shapes.clear();
      return;
    }
    bool known = false;
    for (int i = 0; i < chain_shape_count(); ++i) {
      int g, c;
      chain_shape(i, &g, &c);
      known |= (g == G && c == C);
    }
    if (!known) throw std::invalid_argument("unsupported chain kernel shape");
    if (G * C < m->host->njoints()) throw std::invalid_argument("shape covers fewer joints than the model has");
    auto &v = m->shapes;
    for (auto &s : v)
      if (s.max_batch == max_batch) {
        s.G = G;
        s.C = C;
        return;
      }
    v.push_back({max_batch, G, C});
    std::sort(v.begin(), v.end(), [](const auto &a, const auto &b) { return a.max_batch < b.max_batch; });
  });
}

int sapamd_set_tree_group(sapamd_model *m, int32_t G) {
  return guarded([&] {
    if (!m) throw std::invalid_argument("null model");
    const int nj = m->host->njoints();
    if (G != 0 && ((G != 8 && G != 16 && G != 32) || G < nj))
      throw std::invalid_argument("tree group: 8, 16 or 32 lanes, at least the joint count (0 = default)");
    m->tree_G = G;
  });
}

int sapamd_get_tree_group(const sapamd_model *m, int32_t *G) {
  return guarded([&] {
    if (!m || !G) throw std::invalid_argument("null argument");
    *G = m->tree_lanes();
  });
}

int sapamd_get_chain_shape(const sapamd_model *m, int64_t B, int32_t *G, int32_t *C) {
  return guarded([&] {
    int g, c;
    m->shape_for(B, &g, &c);
    *G = g;
    *C = c;
  });
}

namespace {
// Launch of the fused FK / Jacobian / RNEA kernel family for one model; flags
// carry the outputs and RNEA variants, o the pointers and strides.
void fused_launch(const sapamd_model *m, int64_t B, int flags, const float *q, const float *qd, const float *qdd,
                  const KinOut<float> &o, void *stream, hipEvent_t ev0 = nullptr, hipEvent_t ev1 = nullptr) {
  if (m->effective_variant() == SAPAMD_KERNEL_LANES_PER_JOINT) {
    if (m->chain_kernel()) {
      int G, C;
      m->shape_for(B, &G, &C);
      HIP_OK(launch_chain_fused(m->t32(), m->host->njoints(), B, flags, q, qd, qdd, o, S(stream), G, C, ev0, ev1));
    } else {
      HIP_OK(launch_tree_fused(m->t32(), m->host->njoints(), B, flags, q, qd, qdd, o, S(stream), m->tree_lanes(), ev0,
                               ev1));
    }
  } else {
    if (ev0 || ev1) throw std::runtime_error("kernel timing events need the lanes-per-joint kernels");
    // the generic kernel reads qd / qdd whenever RNEA runs: alias the unused ones to q
    HIP_OK(launch_kinematics<float>(m->t32(), m->host->njoints(), B, flags, q,
                                    (flags & KF_QD0) || !qd ? q : qd, (flags & KF_QDD0) || !qdd ? q : qdd, o,
                                    S(stream)));
  }
}
void check_ld(int64_t ld, int64_t min, const char *what) {
  if (ld != 0 && ld < min) throw std::invalid_argument(std::string(what) + " row stride smaller than a row");
}
}  // namespace

namespace {
void fk_jac_rnea(const sapamd_model *m, int64_t B, const float *q, int64_t q_ld, const float *qd, int64_t qd_ld,
                 const float *qdd, int64_t qdd_ld, float *poses, int64_t pose_env_stride, int64_t pose_link_stride,
                 float *J, float *tau, int64_t tau_ld, void *stream, hipEvent_t ev0 = nullptr,
                 hipEvent_t ev1 = nullptr);
}  // namespace

int sapamd_fk_jac_rnea_batched(const sapamd_model *m, int64_t B, const float *q, const float *qd,
                               const float *qdd, float *poses, float *J, float *tau, void *stream) {
  return sapamd_fk_jac_rnea_strided(m, B, q, 0, qd, 0, qdd, 0, poses, 0, 0, J, tau, 0, stream);
}

int sapamd_fk_jac_rnea_batched_timed(const sapamd_model *m, int64_t B, const float *q, const float *qd,
                                     const float *qdd, float *poses, float *J, float *tau, void *start_event,
                                     void *stop_event, void *stream) {
  return guarded([&] {
    if (!start_event || !stop_event) throw std::invalid_argument("start and stop events must not be null");
    fk_jac_rnea(m, B, q, 0, qd, 0, qdd, 0, poses, 0, 0, J, tau, 0, stream, static_cast<hipEvent_t>(start_event),
                static_cast<hipEvent_t>(stop_event));
  });
}

int sapamd_fk_jac_rnea_strided(const sapamd_model *m, int64_t B, const float *q, int64_t q_ld, const float *qd,
                               int64_t qd_ld, const float *qdd, int64_t qdd_ld, float *poses,
                               int64_t pose_env_stride, int64_t pose_link_stride, float *J, float *tau,
                               int64_t tau_ld, void *stream) {
  return guarded([&] {
    fk_jac_rnea(m, B, q, q_ld, qd, qd_ld, qdd, qdd_ld, poses, pose_env_stride, pose_link_stride, J, tau, tau_ld,
                stream);
  });
}

namespace {
void fk_jac_rnea(const sapamd_model *m, int64_t B, const float *q, int64_t q_ld, const float *qd, int64_t qd_ld,
                 const float *qdd, int64_t qdd_ld, float *poses, int64_t pose_env_stride, int64_t pose_link_stride,
                 float *J, float *tau, int64_t tau_ld, void *stream, hipEvent_t ev0, hipEvent_t ev1) {
  {
    check_batched(m, B);
    if (B == 0) return;
    need(q, "q");
    const int nv = m->nv(), L = m->L();
    check_ld(q_ld, nv, "q");
    check_ld(qd_ld, nv, "qd");
    check_ld(qdd_ld, nv, "qdd");
    check_ld(tau_ld, nv, "tau");
    check_ld(pose_link_stride, 7, "pose link");
    check_ld(pose_env_stride, (int64_t)(L > 0 ? L - 1 : 0) * (pose_link_stride ? pose_link_stride : 7) + 7, "pose");
    int flags = 0;
    KinOut<float> o;
    o.ld_q = q_ld;
    o.ld_qd = qd_ld;
    o.ld_qdd = qdd_ld;
    o.ld_tau = tau_ld;
    o.pose_env = pose_env_stride;
    o.pose_link = pose_link_stride;
    if (poses) { flags |= KF_POSES; o.poses = poses; if (L == 0) throw std::runtime_error("link index out of bound"); }
    if (J) { flags |= KF_JFULL; o.jfull = J; }
    if (tau) { need(qd, "qd"); need(qdd, "qdd"); flags |= KF_RNEA; o.tau = tau; }
    fused_launch(m, B, flags, q, qd, qdd, o, stream, ev0, ev1);
  }
}
}  // namespace

int sapamd_link_state_strided(const sapamd_model *m, int64_t B, const float *q, int64_t q_ld, const float *qd,
                              int64_t qd_ld, float *body, int64_t body_env_stride, int64_t body_link_stride,
                              void *stream) {
  return guarded([&] {
    check_batched(m, B);
    if (B == 0) return;
    need(q, "q");
    need(body, "body");
    const int nv = m->nv(), L = m->L();
    if (L == 0) throw std::runtime_error("link index out of bound");
    if (!tree_group(m->host->njoints()))
      throw std::invalid_argument("link state: trees of 1 to 32 movable joints only");
    check_ld(q_ld, nv, "q");
    check_ld(qd_ld, nv, "qd");
    check_ld(body_link_stride, 13, "body link");
    const int64_t ls = body_link_stride ? body_link_stride : 13;
    check_ld(body_env_stride, (int64_t)(L - 1) * ls + 13, "body");
    const int64_t es = body_env_stride ? body_env_stride : (int64_t)L * ls;
    HIP_OK(launch_link_state(m->t32(), m->host->njoints(), B, q, q_ld ? q_ld : nv, qd, qd_ld ? qd_ld : nv, body, es,
                             ls, S(stream)));
  });
}

int sapamd_random_qpos_batched(const sapamd_model *m, int64_t B, uint64_t seed, float *out, void *stream) {
  return guarded([&] {
    check_batched(m, B);
    if (B == 0) return;
    need(out, "out");
    const int nv = m->nv();
    if (nv > kRandomMaxDofs) throw std::invalid_argument("random configurations: at most 64 dofs");
    RandomLimits lim{};
    if (sapamd_joint_limits(m, lim.lo, lim.hi) != SAPAMD_OK) throw std::runtime_error(sapamd_last_error());
    HIP_OK(launch_random_qpos<float>(B, nv, seed, lim, out, S(stream)));
  });
}

int sapamd_get_random_qpos(sapamd_model *m, uint64_t seed, double *out) {
  return guarded([&] {
    if (!m || !out) throw std::invalid_argument("null argument");
    m->require_tables();
    const int nv = m->nv();
    if (nv > kRandomMaxDofs) throw std::invalid_argument("random configurations: at most 64 dofs");
    RandomLimits lim{};
    if (sapamd_joint_limits(m, lim.lo, lim.hi) != SAPAMD_OK) throw std::runtime_error(sapamd_last_error());
    DeviceGuard g(m->device);
    HIP_OK(launch_random_qpos<double>(1, nv, seed, lim, m->d_buf, m->stream));
    HIP_OK(hipStreamSynchronize(m->stream));
    for (int u = 0; u < nv; ++u) out[u] = m->h_buf[u];
  });
}

int sapamd_passive_force_batched(const sapamd_model *m, int64_t B, const float *q, int64_t q_ld, const float *qd,
                                 int64_t qd_ld, int32_t gravity, int32_t coriolis_and_centrifugal, float *tau,
                                 int64_t tau_ld, void *stream) {
  return guarded([&] {
    check_batched(m, B);
    if (B == 0) return;
    need(q, "q");
    need(tau, "tau");
    const int nv = m->nv();
    check_ld(q_ld, nv, "q");
    check_ld(qd_ld, nv, "qd");
    check_ld(tau_ld, nv, "tau");
    if (coriolis_and_centrifugal) need(qd, "qd");
    int flags = KF_RNEA | KF_QDD0;
    if (!gravity) flags |= KF_NOGRAV;
    if (!coriolis_and_centrifugal) flags |= KF_QD0;
    KinOut<float> o;
    o.tau = tau;
    o.ld_q = q_ld;
    o.ld_qd = qd_ld;
    o.ld_tau = tau_ld;
    fused_launch(m, B, flags, q, coriolis_and_centrifugal ? qd : nullptr, nullptr, o, stream);
  });
}

int sapamd_fk_batched(const sapamd_model *m, int64_t B, const float *q, float *poses, void *stream) {
  need(poses, "poses");
  return sapamd_fk_jac_rnea_batched(m, B, q, nullptr, nullptr, poses, nullptr, nullptr, stream);
}

int sapamd_jacobian_batched(const sapamd_model *m, int64_t B, const float *q, float *J, void *stream) {
  if (!J) { g_last_error = "J must not be null"; return SAPAMD_INVALID_ARGUMENT; }
  return sapamd_fk_jac_rnea_batched(m, B, q, nullptr, nullptr, nullptr, J, nullptr, stream);
}

int sapamd_link_jacobians_batched(const sapamd_model *m, int64_t B, const float *q, float *J_world, float *J_local,
                                  void *stream) {
  return guarded([&] {
    check_batched(m, B);
    if (B == 0) return;
    need(q, "q");
    if (!J_world || !J_local) throw std::invalid_argument("J_world and J_local must both be given");
    KinOut<float> o;
    o.jw_links = J_world;
    o.jl_links = J_local;
    HIP_OK(launch_kinematics<float>(m->t32(), m->host->njoints(), B, KF_JLINKS, q, nullptr, nullptr, o, S(stream)));
  });
}

int sapamd_rnea_batched(const sapamd_model *m, int64_t B, const float *q, const float *qd, const float *qdd,
                        float *tau, void *stream) {
  if (!tau) { g_last_error = "tau must not be null"; return SAPAMD_INVALID_ARGUMENT; }
  return sapamd_fk_jac_rnea_batched(m, B, q, qd, qdd, nullptr, nullptr, tau, stream);
}

int sapamd_crba_batched(const sapamd_model *m, int64_t B, const float *q, float *M, void *stream) {
  return guarded([&] {
    check_batched(m, B);
    if (B == 0) return;
    need(q, "q");
    need(M, "M");
    if (m->effective_variant() == SAPAMD_KERNEL_LANES_PER_JOINT && tree_group(m->host->njoints())) {
      KinOut<float> o;
      o.M = M;
      HIP_OK(launch_tree_fused(m->t32(), m->host->njoints(), B, KF_CRBA, q, nullptr, nullptr, o, S(stream)));
    } else
      HIP_OK(launch_crba<float>(m->t32(), m->host->njoints(), B, q, M, S(stream)));
  });
}

int sapamd_hetero_batched(const sapamd_model *const *models, int32_t nmodels, const int64_t *B,
                          const float *const *q, const float *const *qd, const float *const *qdd,
                          float *const *poses, float *const *J, float *const *tau, float *const *M, void *stream) {
  return guarded([&] {
    if (nmodels < 1 || nmodels > kMaxSegments) throw std::invalid_argument("1 to 8 models per heterogeneous launch");
    HeteroArgs args{};
    args.nseg = 0;
    int64_t blocks = 0;
    int flags = KF_POSES | KF_JFULL | KF_RNEA | KF_CRBA;
    for (int i = 0; i < nmodels; ++i) {
      const sapamd_model *m = models[i];
      check_batched(m, B[i]);
      const int nj = m->host->njoints();
      if (!tree_group(nj)) throw std::runtime_error("heterogeneous launches need trees of 1 to 32 joints");
      const int G = m->tree_lanes();
      if (B[i] == 0) continue;
      need(q[i], "q");
      if (tau && tau[i]) {
        need(qd ? qd[i] : nullptr, "qd");
        need(qdd ? qdd[i] : nullptr, "qdd");
      }
      if (poses && poses[i] && m->L() == 0) throw std::runtime_error("link index out of bound");
      HeteroSegment &s = args.seg[args.nseg++];
      s.tab = m->t32();
      s.G = G;
      s.nj = nj;
      s.B = B[i];
      s.q = q[i];
      s.qd = qd ? qd[i] : nullptr;
      s.qdd = qdd ? qdd[i] : nullptr;
      s.poses = poses ? poses[i] : nullptr;
      s.J = J ? J[i] : nullptr;
      s.tau = tau ? tau[i] : nullptr;
      s.M = M ? M[i] : nullptr;
    }
    // blocks of the costliest trees first (most lanes per environment, then
    // most joints): their wavefronts bound the launch, so they start first
    std::stable_sort(args.seg, args.seg + args.nseg, [](const HeteroSegment &x, const HeteroSegment &y) {
      return x.G != y.G ? x.G > y.G : x.nj > y.nj;
    });
    for (int i = 0; i < args.nseg; ++i) {
      args.seg[i].block0 = blocks;
      blocks += (args.seg[i].B * args.seg[i].G + 255) / 256;
    }
    HIP_OK(launch_hetero(args, blocks, flags, S(stream)));
  });
}


// ---------------------------------------------------------------- CLIK schedule
// Library-wide (not per model), so a problem gives the same bits in a
// per-model and in a mixed-tree launch.
namespace {
std::mutex g_clik_mu;
ClikSchedule g_clik_schedule;
ClikSchedule clik_schedule() {
  std::lock_guard<std::mutex> lock(g_clik_mu);
  return g_clik_schedule;
}
}  // namespace

int sapamd_set_clik_schedule(int32_t cap, int32_t g1, int32_t g2) {
  return guarded([&] {
    ClikSchedule sch;
    sch.cap = cap;
    sch.g1 = g1;
    sch.g2 = g2;
    if (cap == 0 && (g2 != 4 && g2 != 8 && g2 != 16)) sch.g2 = 8;
    if (!clik_schedule_valid(sch)) throw std::invalid_argument("CLIK schedule: cap >= 0, g1 and g2 in {4, 8, 16}");
    std::lock_guard<std::mutex> lock(g_clik_mu);
    g_clik_schedule = sch;
  });
}

int sapamd_get_clik_schedule(int32_t *cap, int32_t *g1, int32_t *g2) {
  return guarded([&] {
    const ClikSchedule sch = clik_schedule();
    if (cap) *cap = sch.cap;
    if (g1) *g1 = sch.g1;
    if (g2) *g2 = sch.g2;
  });
}

// ---------------------------------------------------------------- allocation-free CLIK
// Scratch memory of sapamd_clik_batched_noalloc: the four queue counters of a
// CLIK launch, then one hand-over record per problem, for at most
// kNoAllocClikProblems problems (about 11 MB per GPU, held until the process
// ends). Host code never passes `scratch` itself to a kernel or copy: it is a
// host-side stand-in; each device's copy is reached through the address that
// hipGetSymbolAddress returns with that device current (noalloc_scratch).
constexpr int64_t kNoAllocClikProblems = 65536;
constexpr size_t kNoAllocCounterBytes = 256;
static __device__ __align__(256) char scratch[kNoAllocCounterBytes + sizeof(ClikCont) * kNoAllocClikProblems];

namespace {
constexpr int kMaxDevices = 64;

// This device's copy of `scratch`, looked up once per device.
char *noalloc_scratch(int device) {
  static std::mutex mu;
  static char *ptr[kMaxDevices] = {};
  if (device < 0 || device >= kMaxDevices) throw std::runtime_error("device index out of range");
  std::lock_guard<std::mutex> lock(mu);
  if (!ptr[device]) {
    DeviceGuard g(device);
    void *p = nullptr;
    HIP_OK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    ptr[device] = static_cast<char *>(p);
  }
  return ptr[device];
}

// Compute units of a device (queried once per device).
int device_cus(int device) {
  static std::mutex mu;
  static int cus[kMaxDevices] = {};
  if (device < 0 || device >= kMaxDevices) return 256;
  std::lock_guard<std::mutex> lock(mu);
  if (!cus[device]) {
    int n = 0;
    if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess || n <= 0) {
      (void)hipGetLastError();
      n = 256;
    }
    cus[device] = n;
  }
  return cus[device];
}

}  // namespace

int sapamd_clik_hetero_batched(const sapamd_model *const *models, int32_t nmodels, const int64_t *B,
                               const int32_t *links, const float *const *targets, const float *const *q0,
                               const int32_t *const *masks, double eps, int32_t max_iterations, double dt,
                               double damp, float *const *q_out, int32_t *const *success, float *const *err,
                               int32_t *const *iterations, void *stream) {
  return guarded([&] {
    if (nmodels < 1 || nmodels > kMaxSegments) throw std::invalid_argument("1 to 8 models per heterogeneous launch");
    need(B, "B");
    need(links, "links");
    need(targets, "targets");
    need(q_out, "q_out");
    need(success, "success");
    need(err, "err");
    HeteroClikArgs A{};
    A.nseg = 0;
    A.total = 0;
    int max_chain = 1;
    for (int i = 0; i < nmodels; ++i) {
      const sapamd_model *m = models[i];
      check_batched(m, B[i]);
      m->check_link(links[i]);
      if (B[i] == 0) continue;
      need(targets[i], "target");
      need(q_out[i], "q_out");
      need(success[i], "success");
      need(err[i], "err");
      ClikSegment &g = A.seg[A.nseg++];
      g.tab = m->t32();
      g.P.link = links[i];
      g.P.chain = m->d_chain + m->chain_off[links[i]];
      g.P.nchain = m->chain_len[links[i]];
      g.P.eps = eps;
      g.P.dt = dt;
      g.P.damp = damp;
      g.P.max_iter = max_iterations;
      g.P.mask = masks ? masks[i] : nullptr;
      g.P.mask_ld = 0;
      g.P.counter = nullptr;
      g.P.persistent_waves = 0;
      g.target = targets[i];
      g.q0 = q0 ? q0[i] : nullptr;
      g.q_out = q_out[i];
      g.success = success[i];
      g.err = err[i];
      g.iters = iterations ? iterations[i] : nullptr;
      g.B = B[i];
      g.first = A.total;
      A.total += B[i];
      max_chain = std::max(max_chain, g.P.nchain);
    }
    if (A.total == 0) return;
    // the queue counter (and hand-over records) of the first model:
    // overlapping launches on it race
    A.counter = models[0]->d_counter;
    const int cus = device_cus(models[0]->device);
    bool lanes = clik_lanes_group(max_chain) != 0 && A.total <= INT32_MAX;
    for (int i = 0; i < nmodels; ++i) lanes = lanes && models[i]->variant != SAPAMD_KERNEL_THREAD_PER_ENV;
    if (lanes) {
      models[0]->clik_workspace(A, A.total, S(stream));
      HIP_OK(launch_clik_lanes(A, max_chain, cus, clik_schedule(), S(stream)));
      return;
    }
    HIP_OK(launch_fill_i32(A.counter, 0, 1, S(stream)));
    const int64_t waves = std::min<int64_t>(4 * cus, (A.total + 63) / 64);
    HIP_OK(launch_clik_hetero(A, max_chain, (int)waves, S(stream)));
  });
}

int sapamd_coriolis_batched(const sapamd_model *m, int64_t B, const float *q, const float *qd, float *C,
                            void *stream) {
  return guarded([&] {
    check_batched(m, B);
    if (B == 0) return;
    need(q, "q");
    need(qd, "qd");
    need(C, "C");
    HIP_OK(launch_coriolis<float>(m->t32(), m->host->njoints(), B, q, qd, C, S(stream)));
  });
}

namespace {
// Trees of <= 32 joints under the lanes-per-joint kernels factor M in LDS (FP32
// or, past the condition limit, FP64); the rest take the thread-per-env ABA /
// FP64 factorization.
bool lanes_dynamics(const sapamd_model *m) {
  return m->effective_variant() == SAPAMD_KERNEL_LANES_PER_JOINT && tree_dyn_group(m->host->njoints()) != 0;
}
}  // namespace

int sapamd_aba_batched(const sapamd_model *m, int64_t B, const float *q, const float *qd, const float *tau,
                       float *qdd, void *stream) {
  return guarded([&] {
    check_batched(m, B);
    if (B == 0) return;
    need(q, "q");
    need(qd, "qd");
    need(tau, "tau");
    need(qdd, "qdd");
    if (lanes_dynamics(m))
      HIP_OK(launch_tree_step(m->t32(), m->fd64() ? m->t64() : nullptr, m->host->njoints(), B, 0, 0.f,
                              const_cast<float *>(q), const_cast<float *>(qd), tau, m->nv(), qdd, nullptr, 0,
                              S(stream), m->dyn_rows));
    else if (m->fd64())
      HIP_OK(launch_fd_ldlt(m->t64(), m->host->njoints(), B, q, qd, tau, qdd, S(stream)));
    else
      HIP_OK(launch_aba<float>(m->t32(), m->host->njoints(), B, q, qd, tau, qdd, S(stream)));
  });
}

namespace {
// Batched CLIK over nstart x B problems (start-major); solved != nullptr turns
// on the multi-start skip. With ws != nullptr the queue counters and hand-over
// records are taken from ws (the allocation-free entry point's scratch: four
// counters, then records from byte kNoAllocCounterBytes on) instead of the model.
void clik_launch(const sapamd_model *m, int64_t B, int64_t nstart, int32_t link, const float *target,
                 const float *q0, const int32_t *mask, double eps, int32_t max_iterations, double dt, double damp,
                 float *q_out, int32_t *success, float *err, int32_t *iterations, int32_t *solved, void *stream,
                 char *ws = nullptr) {
  int32_t *counters = ws ? reinterpret_cast<int32_t *>(ws) : m->d_counter;
  ClikParams P;
  P.link = link;
  P.chain = m->d_chain + m->chain_off[link];
  P.nchain = m->chain_len[link];
  P.eps = eps;
  P.dt = dt;
  P.damp = damp;
  P.max_iter = max_iterations;
  P.mask = mask;
  P.mask_ld = 0;
  P.counter = counters;
  P.ntargets = solved ? B : 0;
  P.solved = solved;
  const int64_t total = B * nstart;
  const int dev_cus = device_cus(m->device);
  if (clik_lanes_group(P.nchain) && total <= INT32_MAX && m->variant != SAPAMD_KERNEL_THREAD_PER_ENV) {
    // one segment of the lanes-per-joint CLIK (problems taken in index order,
    // so multi-start problems run start-major)
    HeteroClikArgs A{};
    ClikSegment &g = A.seg[0];
    g.tab = m->t32();
    g.P = P;
    g.target = target;
    g.q0 = q0;
    g.q_out = q_out;
    g.err = err;
    g.success = success;
    g.iters = iterations;
    g.B = total;
    g.first = 0;
    A.nseg = 1;
    A.total = total;
    if (ws) {
      A.counter = counters;
      A.cont_count = counters + 1;
      A.cont_queue = counters + 2;
      A.cont_easy = counters + 3;
      A.cont = reinterpret_cast<ClikCont *>(ws + kNoAllocCounterBytes);
      HIP_OK(launch_fill_i32(counters, 0, 4, S(stream)));
    } else {
      m->clik_workspace(A, total, S(stream));
    }
    if (solved) HIP_OK(launch_fill_i32(solved, INT32_MAX, B, S(stream)));
    HIP_OK(launch_clik_lanes(A, P.nchain, dev_cus, clik_schedule(), S(stream)));
    return;
  }
  const int64_t waves = (total + 63) / 64;
  // Persistent launch with a work queue once there are more problems than
  // resident lanes (and always for multi-start, so that starts are taken in
  // order): finished lanes pull the next problem. The queue counter belongs to
  // the model, so overlapping batched CLIK calls on one model race.
  P.persistent_waves = (waves > 4 * dev_cus || solved) ? (int)std::min<int64_t>(waves, 4 * dev_cus) : 0;
  if (P.persistent_waves) HIP_OK(launch_fill_i32(counters, 0, 1, S(stream)));
  else P.counter = nullptr;
  if (solved) HIP_OK(launch_fill_i32(solved, INT32_MAX, B, S(stream)));
  HIP_OK(launch_clik<float>(m->t32(), m->host->njoints(), total, P, target, q0, q_out, success, err, iterations,
                            S(stream)));
}
}  // namespace

int sapamd_clik_batched(const sapamd_model *m, int64_t B, int32_t link, const float *target, const float *q0,
                        const int32_t *mask, double eps, int32_t max_iterations, double dt, double damp,
                        float *q_out, int32_t *success, float *err, int32_t *iterations, void *stream) {
  return guarded([&] {
    check_batched(m, B);
    m->check_link(link);
    if (B == 0) return;
    need(target, "target");
    need(q_out, "q_out");
    need(success, "success");
    need(err, "err");
    clik_launch(m, B, 1, link, target, q0, mask, eps, max_iterations, dt, damp, q_out, success, err, iterations,
                nullptr, stream);
  });
}

int sapamd_clik_batched_noalloc(const sapamd_model *m, int64_t B, int32_t link, const float *target,
                                const float *q0, const int32_t *mask, double eps, int32_t max_iterations, double dt,
                                double damp, float *q_out, int32_t *success, float *err, int32_t *iterations,
                                void *stream) {
  return guarded([&] {
    check_batched(m, B);
    m->check_link(link);
    if (B > kNoAllocClikProblems)
      throw std::invalid_argument("sapamd_clik_batched_noalloc: at most 65536 problems per call");
    if (B == 0) return;
    need(target, "target");
    need(q_out, "q_out");
    need(success, "success");
    need(err, "err");
    char *ws = noalloc_scratch(m->device);
    clik_launch(m, B, 1, link, target, q0, mask, eps, max_iterations, dt, damp, q_out, success, err, iterations,
                nullptr, stream, ws);
  });
}

int sapamd_clik_multistart_batched(const sapamd_model *m, int64_t B, int32_t nstart, int32_t link,
                                   const float *target, const float *q0, const int32_t *mask, double eps,
                                   int32_t max_iterations, double dt, double damp, float *q_out, int32_t *success,
                                   float *err, int32_t *iterations, int32_t *solved, void *stream) {
  return guarded([&] {
    check_batched(m, B);
    m->check_link(link);
    if (nstart < 1) throw std::invalid_argument("at least one start per target");
    if (B == 0) return;
    for (const void *p : {(const void *)target, (const void *)q0, (const void *)q_out, (const void *)success,
                          (const void *)err, (const void *)solved})
      need(p, "multi-start buffer");
    clik_launch(m, B, nstart, link, target, q0, mask, eps, max_iterations, dt, damp, q_out, success, err, iterations,
                solved, stream);
  });
}

namespace {
// Chain kernels need one user link per joint for pose wrenches; otherwise the
// thread-per-environment kernel runs.
bool links_one_per_joint(const sapamd_model *m) {
  const auto &lf = m->host->link_frames();
  std::vector<int> cnt(m->host->njoints(), 0);
  int root = 0;
  for (int f : lf) {
    int j = m->host->frames()[f].joint;
    if (j < 0) ++root;
    else if (++cnt[j] > 1) return false;
  }
  return root <= 1;
}
void run_vjp(const sapamd_model *m, int64_t B, const float *q, const float *qd, const float *qdd,
             VjpArgs<float> a, bool poses, void *stream) {
  const bool lanes = m->effective_variant() == SAPAMD_KERNEL_LANES_PER_JOINT && (!poses || links_one_per_joint(m));
  if (lanes && m->chain_kernel()) {
    int G, C;
    m->shape_for(B, &G, &C);
    HIP_OK(launch_chain_vjp(m->t32(), m->host->njoints(), B, q, qd, qdd, a, S(stream), G, C));
  } else if (lanes && tree_group(m->host->njoints())) {
    HIP_OK(launch_tree_vjp(m->t32(), m->host->njoints(), B, q, qd, qdd, a, S(stream)));
  } else {
    HIP_OK(launch_kin_vjp<float>(m->t32(), m->host->njoints(), B, q, qd, qdd, a, S(stream)));
  }
}
}  // namespace

int sapamd_kin_vjp_batched(const sapamd_model *m, int64_t B, const float *q, const float *qd, const float *qdd,
                           const float *gposes, const float *gtau, float *dq, float *dqd, float *dqdd, void *stream) {
  return guarded([&] {
    check_batched(m, B);
    if (B == 0) return;
    need(q, "q");
    need(dq, "dq");
    if (!gposes && !gtau) throw std::invalid_argument("give gposes and/or gtau");
    if (gtau) {
      need(qd, "qd");
      need(qdd, "qdd");
      need(dqd, "dqd");
      need(dqdd, "dqdd");
    }
    if (gposes && m->L() == 0) throw std::runtime_error("link index out of bound");
    VjpArgs<float> a;
    a.gposes = gposes;
    a.gtau = gtau;
    a.dq = dq;
    a.dqd = gtau ? dqd : nullptr;
    a.dqdd = gtau ? dqdd : nullptr;
    run_vjp(m, B, q, qd, qdd, a, gposes != nullptr, stream);
  });
}

int sapamd_pose_torque_loss_batched(const sapamd_model *m, int64_t B, const float *q, const float *qd,
                                    const float *qdd, const float *tau_star, int32_t ee_link, const float *p_target,
                                    float *loss, float *tau, float *dq, float *dqd, float *dqdd, void *stream) {
  return guarded([&] {
    check_batched(m, B);
    if (p_target) m->check_link(ee_link);
    if (B == 0) return;
    need(q, "q");
    need(loss, "loss");
    need(dq, "dq");
    if (!tau_star && !p_target) throw std::invalid_argument("give tau_star and/or p_target");
    if (tau_star) {
      need(qd, "qd");
      need(qdd, "qdd");
      need(dqd, "dqd");
      need(dqdd, "dqdd");
    }
    VjpArgs<float> a;
    a.tau_star = tau_star;
    a.p_target = p_target;
    a.ee_link = ee_link;
    a.loss = loss;
    a.tau = tau;
    a.dq = dq;
    a.dqd = tau_star ? dqd : nullptr;
    a.dqdd = tau_star ? dqdd : nullptr;
    run_vjp(m, B, q, qd, qdd, a, p_target != nullptr, stream);
  });
}

namespace {
// Forward dynamics of one step: the lanes-per-joint FP32 factorization for
// trees of <= 32 joints (lanes-per-joint variant), the thread-per-environment
// kernels otherwise (FP32 ABA, or FP64 CRBA + LTDL for ill-conditioned models).

void step_launch(const sapamd_model *m, int64_t B, float dt, float *q, float *qd, const float *tau,
                 int64_t tau_stride, float *save, int64_t save_stride, void *stream) {
  if (lanes_dynamics(m))
    HIP_OK(launch_tree_step(m->t32(), m->fd64() ? m->t64() : nullptr, m->host->njoints(), B, 1, dt, q, qd, tau,
                            tau_stride, nullptr, save,
                            save_stride, S(stream), m->dyn_rows));
  else
    HIP_OK(launch_rollout_step<float>(m->t32(), m->fd64() ? m->t64() : nullptr, m->host->njoints(), B, dt, q, qd,
                                      tau, tau_stride, save, save_stride, S(stream)));
}
void step_vjp_launch(const sapamd_model *m, int64_t B, float dt, const float *state, int64_t state_stride,
                     const float *tau, int64_t tau_stride, float *aq, float *aqd, float *gtau, void *stream) {
  if (lanes_dynamics(m))
    HIP_OK(launch_tree_step_vjp(m->t32(), m->fd64() ? m->t64() : nullptr, m->host->njoints(), B, dt, state,
                                state_stride, tau, tau_stride, aq, aqd,
                                gtau, S(stream), m->dyn_rows));
  else
    HIP_OK(launch_rollout_step_vjp<float>(m->t32(), m->fd64() ? m->t64() : nullptr, m->host->njoints(), B, dt, state,
                                          state_stride, tau, tau_stride, aq, aqd, gtau, S(stream)));
}
}  // namespace

int sapamd_rollout_step(const sapamd_model *m, int64_t B, float dt, float *q, float *qd, const float *tau_t,
                        int64_t tau_stride, void *stream) {
  return guarded([&] {
    check_batched(m, B);
    if (B == 0) return;
    need(q, "q");
    need(qd, "qd");
    need(tau_t, "tau_t");
    step_launch(m, B, dt, q, qd, tau_t, tau_stride, nullptr, 0, stream);
  });
}

int sapamd_rollout_forward(const sapamd_model *m, int64_t B, int32_t T, int32_t K, float dt, float *q, float *qd,
                           const float *tau, float *ckpt, void *stream) {
  return guarded([&] {
    check_batched(m, B);
    if (T < 0 || K < 1) throw std::invalid_argument("T >= 0 and K >= 1 required");
    if (B == 0 || T == 0) return;
    need(q, "q");
    need(qd, "qd");
    need(tau, "tau");
    if (lanes_dynamics(m) && !m->per_step_rollouts) {
      HIP_OK(launch_tree_rollout(m->t32(), m->fd64() ? m->t64() : nullptr, m->host->njoints(), B, T, K, dt, q, qd,
                                 tau, ckpt, S(stream), m->dyn_rows));
      return;
    }
    const int nv = m->nv();
    const int64_t S_ = (T + K - 1) / K;
    for (int t = 0; t < T; ++t) {
      float *save = (ckpt && t % K == 0) ? ckpt + (int64_t)(t / K) * 2 * nv : nullptr;
      step_launch(m, B, dt, q, qd, tau + (int64_t)t * nv, (int64_t)T * nv, save, S_ * 2 * nv, stream);
    }
  });
}

int sapamd_rollout_backward(const sapamd_model *m, int64_t B, int32_t T, int32_t K, float dt, const float *tau,
                            const float *ckpt, float *seg, float *q_tmp, float *qd_tmp, float *aq, float *aqd,
                            float *gtau, void *stream) {
  return guarded([&] {
    check_batched(m, B);
    if (T < 0 || K < 1) throw std::invalid_argument("T >= 0 and K >= 1 required");
    if (B == 0 || T == 0) return;
    for (const void *p : {(const void *)tau, (const void *)ckpt, (const void *)seg, (const void *)q_tmp,
                          (const void *)qd_tmp, (const void *)aq, (const void *)aqd, (const void *)gtau})
      need(p, "rollout buffer");
    if (lanes_dynamics(m) && !m->per_step_rollouts) {
      HIP_OK(launch_tree_rollout_vjp(m->t32(), m->fd64() ? m->t64() : nullptr, m->host->njoints(), B, T, K, dt, tau,
                                     ckpt, seg, aq, aqd, gtau, S(stream), m->dyn_rows));
      return;
    }
    const int nv = m->nv();
    const int64_t S_ = (T + K - 1) / K;
    for (int64_t s = S_ - 1; s >= 0; --s) {
      const int t0 = (int)(s * K), t1 = std::min(T, t0 + K);
      HIP_OK(launch_restore_state(B, nv, ckpt + s * 2 * nv, S_ * 2 * nv, q_tmp, qd_tmp, S(stream)));
      for (int t = t0; t < t1; ++t)  // recompute the segment, keeping every pre-step state
        step_launch(m, B, dt, q_tmp, qd_tmp, tau + (int64_t)t * nv, (int64_t)T * nv,
                    seg + (int64_t)(t - t0) * 2 * nv, (int64_t)K * 2 * nv, stream);
      for (int t = t1 - 1; t >= t0; --t)
        step_vjp_launch(m, B, dt, seg + (int64_t)(t - t0) * 2 * nv, (int64_t)K * 2 * nv, tau + (int64_t)t * nv,
                        (int64_t)T * nv, aq, aqd, gtau + (int64_t)t * nv, stream);
    }
  });
}

namespace {
// One segment per model for a mixed-tree rollout launch; models with no
// environments are left out.
RolloutMixArgs rollout_mix(const sapamd_model *const *models, int32_t nmodels, const int64_t *B, int32_t T,
                           int32_t K, float dt) {
  if (nmodels < 1 || nmodels > kMaxSegments) throw std::invalid_argument("1 to 8 models per mixed-tree launch");
  if (!models || !B) throw std::invalid_argument("models and B must not be null");
  if (T < 0 || K < 1) throw std::invalid_argument("T >= 0 and K >= 1 required");
  RolloutMixArgs A{};
  A.T = T;
  A.K = K;
  A.dt = dt;
  for (int i = 0; i < nmodels; ++i) {
    const sapamd_model *m = models[i];
    check_batched(m, B[i]);
    if (!lanes_dynamics(m))
      throw std::runtime_error("mixed-tree rollouts need trees of 1 to 32 joints under the lanes-per-joint kernels");
    if (m->device != models[0]->device) throw std::invalid_argument("mixed-tree rollouts need models on one GPU");
    RolloutSegment &s = A.seg[i];
    s.fp64 = m->fd64() ? 1 : 0;
    s.tab = s.fp64 ? static_cast<const void *>(m->t64()) : static_cast<const void *>(m->t32());
    s.nj = m->host->njoints();
    s.rows = m->dyn_rows;
    s.B = B[i];
  }
  A.nseg = nmodels;
  return A;
}
}  // namespace

int sapamd_rollout_forward_hetero(const sapamd_model *const *models, int32_t nmodels, const int64_t *B, int32_t T,
                                  int32_t K, float dt, float *const *q, float *const *qd, const float *const *tau,
                                  float *const *ckpt, void *stream) {
  return guarded([&] {
    RolloutMixArgs A = rollout_mix(models, nmodels, B, T, K, dt);
    if (T == 0) return;
    need(q, "q");
    need(qd, "qd");
    need(tau, "tau");
    for (int i = 0; i < nmodels; ++i) {
      if (B[i] == 0) continue;
      need(q[i], "q");
      need(qd[i], "qd");
      need(tau[i], "tau");
      RolloutSegment &s = A.seg[i];
      s.q = q[i];
      s.qd = qd[i];
      s.tau = tau[i];
      s.ckpt_out = ckpt ? ckpt[i] : nullptr;
    }
    HIP_OK(launch_rollout_mixed(A, false, S(stream)));
  });
}

int sapamd_rollout_backward_hetero(const sapamd_model *const *models, int32_t nmodels, const int64_t *B, int32_t T,
                                   int32_t K, float dt, const float *const *tau, const float *const *ckpt,
                                   float *const *seg, float *const *aq, float *const *aqd, float *const *gtau,
                                   void *stream) {
  return guarded([&] {
    RolloutMixArgs A = rollout_mix(models, nmodels, B, T, K, dt);
    if (T == 0) return;
    for (const void *p : {(const void *)tau, (const void *)ckpt, (const void *)seg, (const void *)aq,
                          (const void *)aqd, (const void *)gtau})
      need(p, "rollout buffer array");
    for (int i = 0; i < nmodels; ++i) {
      if (B[i] == 0) continue;
      for (const void *p : {(const void *)tau[i], (const void *)ckpt[i], (const void *)seg[i], (const void *)aq[i],
                            (const void *)aqd[i], (const void *)gtau[i]})
        need(p, "rollout buffer");
      RolloutSegment &s = A.seg[i];
      s.tau = tau[i];
      s.ckpt = ckpt[i];
      s.seg = seg[i];
      s.aq = aq[i];
      s.aqd = aqd[i];
      s.gtau = gtau[i];
    }
    HIP_OK(launch_rollout_mixed(A, true, S(stream)));
  });
}

int sapamd_set_rollout_launch(sapamd_model *m, int32_t mode) {
  return guarded([&] {
    if (!m) throw std::invalid_argument("null model");
    if (mode != SAPAMD_ROLLOUT_FUSED && mode != SAPAMD_ROLLOUT_PER_STEP)
      throw std::invalid_argument("unknown rollout launch mode");
    m->per_step_rollouts = mode == SAPAMD_ROLLOUT_PER_STEP;
  });
}

int sapamd_set_dynamics_precision(sapamd_model *m, int32_t mode) {
  return guarded([&] {
    if (!m) throw std::invalid_argument("null model");
    if (mode < SAPAMD_DYNAMICS_AUTO || mode > SAPAMD_DYNAMICS_FP64)
      throw std::invalid_argument("unknown dynamics precision");
    m->dyn_precision = mode;
  });
}

int sapamd_get_dynamics_precision(const sapamd_model *m, int32_t *mode, double *cond_estimate) {
  return guarded([&] {
    if (!m) throw std::invalid_argument("null model");
    if (mode) *mode = m->fd64() ? SAPAMD_DYNAMICS_FP64 : SAPAMD_DYNAMICS_FP32;
    if (cond_estimate) *cond_estimate = m->cond_est;
  });
}

}  // extern "C"
