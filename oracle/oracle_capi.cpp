// ORACLE — test infrastructure only (see rbd_oracle.hpp). C ABI over the FP64
// restatement so the Python tests and the CPU-baseline leg of bench.py can call
// it through ctypes. Inputs/outputs are in the user (SAPIEN) dof order and link
// order, like the reference's public methods (pinocchio_model.cpp:146-300).
#include "rbd_oracle.hpp"

#include <chrono>
#include <cstring>
#include <thread>

using namespace oracle;

namespace {

template <class S> void pose7(const SE3<S> &M, S *out) {
  for (int i = 0; i < 3; ++i) out[i] = M.p(i, 0);
  quat_from_matrix(M.R, out + 3);
}

SE3<double> se3_from_pose7(const double *pq) {
  // pose given as p(3) + quaternion wxyz(4), converted as
  // Eigen::Quaterniond(w, x, y, z).toRotationMatrix() does (pinocchio_model.cpp:255):
  // the unit-quaternion formula applied as is, without normalizing q.
  const double w = pq[3], x = pq[4], y = pq[5], z = pq[6];
  SE3<double> M;
  M.R(0, 0) = 1 - 2 * (y * y + z * z); M.R(0, 1) = 2 * (x * y - w * z); M.R(0, 2) = 2 * (x * z + w * y);
  M.R(1, 0) = 2 * (x * y + w * z); M.R(1, 1) = 1 - 2 * (x * x + z * z); M.R(1, 2) = 2 * (y * z - w * x);
  M.R(2, 0) = 2 * (x * z - w * y); M.R(2, 1) = 2 * (y * z + w * x); M.R(2, 2) = 1 - 2 * (x * x + y * y);
  for (int i = 0; i < 3; ++i) M.p(i, 0) = pq[i];
  return M;
}

template <class F> void parallel_for(int64_t B, int threads, F fn) {
  if (threads <= 1 || B < 2) {
    for (int64_t b = 0; b < B; ++b) fn(b);
    return;
  }
  std::vector<std::thread> pool;
  int64_t chunk = (B + threads - 1) / threads;
  for (int t = 0; t < threads; ++t) {
    int64_t lo = t * chunk, hi = std::min(B, lo + chunk);
    if (lo >= hi) break;
    pool.emplace_back([=] {
      for (int64_t b = lo; b < hi; ++b) fn(b);
    });
  }
  for (auto &th : pool) th.join();
}

}  // namespace

extern "C" {

void *oracle_create(int nj, const int *parent, const int *jtype, const double *axis,
                    const double *jplace, const double *inertia, const int *vuser, int nlinks,
                    const int *link_joint, const double *link_place, const double *gravity) {
  auto *m = new Model;
  m->nj = nj;
  m->parent.assign(parent, parent + nj);
  m->jtype.assign(jtype, jtype + nj);
  m->vuser.assign(vuser, vuser + nj);
  m->axis.resize(nj);
  m->jplace.resize(nj);
  m->inertia.resize(nj);
  for (int j = 0; j < nj; ++j) {
    std::memcpy(m->axis[j].data(), axis + 3 * j, 3 * sizeof(double));
    std::memcpy(m->jplace[j].data(), jplace + 12 * j, 12 * sizeof(double));
    std::memcpy(m->inertia[j].data(), inertia + 10 * j, 10 * sizeof(double));
  }
  m->nlinks = nlinks;
  m->link_joint.assign(link_joint, link_joint + nlinks);
  m->link_place.resize(nlinks);
  for (int l = 0; l < nlinks; ++l)
    std::memcpy(m->link_place[l].data(), link_place + 12 * l, 12 * sizeof(double));
  for (int i = 0; i < 3; ++i) m->gravity[i] = gravity[i];
  return m;
}

void oracle_destroy(void *h) { delete static_cast<Model *>(h); }

// Link poses (B, L, 7) = p xyz | q wxyz (pinocchio_model.cpp:150-161).
void oracle_fk_poses(void *h, int64_t B, const double *q, double *poses) {
  const Model &m = *static_cast<Model *>(h);
  int n = m.nj, L = m.nlinks;
  for (int64_t b = 0; b < B; ++b) {
    std::vector<SE3<double>> oMi;
    forward_kinematics(m, q + b * n, oMi);
    for (int l = 0; l < L; ++l) pose7(link_pose(m, oMi, l), poses + (b * L + l) * 7);
  }
}

// Link 4x4-free outputs: rotation matrices (B, L, 3, 3) and translations (B, L, 3).
void oracle_fk_mats(void *h, int64_t B, const double *q, double *R, double *p) {
  const Model &m = *static_cast<Model *>(h);
  int n = m.nj, L = m.nlinks;
  for (int64_t b = 0; b < B; ++b) {
    std::vector<SE3<double>> oMi;
    forward_kinematics(m, q + b * n, oMi);
    for (int l = 0; l < L; ++l) {
      SE3<double> M = link_pose(m, oMi, l);
      for (int i = 0; i < 3; ++i) {
        p[(b * L + l) * 3 + i] = M.p(i, 0);
        for (int j = 0; j < 3; ++j) R[((b * L + l) * 3 + i) * 3 + j] = M.R(i, j);
      }
    }
  }
}

// Full WORLD Jacobian (B, 6, nv), user column order (computeJointJacobians).
void oracle_jacobian_world(void *h, int64_t B, const double *q, double *J) {
  const Model &m = *static_cast<Model *>(h);
  int n = m.nj;
  for (int64_t b = 0; b < B; ++b) {
    std::vector<SE3<double>> oMi;
    std::vector<V6<double>> cols;
    forward_kinematics(m, q + b * n, oMi);
    world_jacobian(m, oMi, cols);
    for (int j = 0; j < n; ++j)
      for (int r = 0; r < 6; ++r) J[(b * 6 + r) * n + m.vuser[j]] = cols[j](r, 0);
  }
}

// get_link_jacobian(link, local) after compute_full_jacobian(q): (B, 6, nv).
void oracle_link_jacobian(void *h, int64_t B, const double *q, int link, int local, double *J) {
  const Model &m = *static_cast<Model *>(h);
  int n = m.nj;
  for (int64_t b = 0; b < B; ++b) {
    std::vector<SE3<double>> oMi;
    std::vector<V6<double>> cols;
    forward_kinematics(m, q + b * n, oMi);
    world_jacobian(m, oMi, cols);
    link_jacobian(m, oMi, cols, link, local != 0, J + b * 6 * n);
  }
}

void oracle_rnea(void *h, int64_t B, const double *q, const double *qd, const double *qdd,
                 double *tau) {
  const Model &m = *static_cast<Model *>(h);
  int n = m.nj;
  for (int64_t b = 0; b < B; ++b) rnea(m, q + b * n, qd + b * n, qdd + b * n, tau + b * n);
}

void oracle_crba(void *h, int64_t B, const double *q, double *M) {
  const Model &m = *static_cast<Model *>(h);
  int n = m.nj;
  for (int64_t b = 0; b < B; ++b) crba(m, q + b * n, M + b * n * n);
}

void oracle_aba(void *h, int64_t B, const double *q, const double *qd, const double *tau,
                double *qdd) {
  const Model &m = *static_cast<Model *>(h);
  int n = m.nj;
  for (int64_t b = 0; b < B; ++b) aba(m, q + b * n, qd + b * n, tau + b * n, qdd + b * n);
}

// Christoffel-consistent Coriolis matrix from exact (dual-number) derivatives of
// the CRBA mass matrix: C_ij = sum_k 1/2 (dM_ij/dq_k + dM_ik/dq_j - dM_jk/dq_i) qd_k.
void oracle_coriolis(void *h, int64_t B, const double *q, const double *qd, double *C) {
  const Model &m = *static_cast<Model *>(h);
  int n = m.nj;
  std::vector<Dual> qq(n), MM(n * n);
  std::vector<double> dM(n * n * n);  // dM[k][i][j]
  for (int64_t b = 0; b < B; ++b) {
    for (int k = 0; k < n; ++k) {
      for (int i = 0; i < n; ++i) qq[i] = Dual(q[b * n + i], i == k ? 1.0 : 0.0);
      crba(m, qq.data(), MM.data());
      for (int i = 0; i < n * n; ++i) dM[k * n * n + i] = MM[i].d;
    }
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < n; ++j) {
        double acc = 0;
        for (int k = 0; k < n; ++k)
          acc += 0.5 * (dM[k * n * n + i * n + j] + dM[j * n * n + i * n + k] -
                        dM[i * n * n + j * n + k]) *
                 qd[b * n + k];
        C[b * n * n + i * n + j] = acc;
      }
  }
}

// Exact partial derivatives of RNEA by forward-mode dual numbers.
// d*[b][i][k] = d tau_i / d x_k for x in {q, qd, qdd}.
void oracle_rnea_derivs(void *h, int64_t B, const double *q, const double *qd, const double *qdd,
                        double *dq, double *dqd, double *dqdd) {
  const Model &m = *static_cast<Model *>(h);
  int n = m.nj;
  std::vector<Dual> a(n), b_(n), c(n), t(n);
  for (int64_t b = 0; b < B; ++b) {
    for (int which = 0; which < 3; ++which) {
      double *out = which == 0 ? dq : which == 1 ? dqd : dqdd;
      for (int k = 0; k < n; ++k) {
        for (int i = 0; i < n; ++i) {
          a[i] = Dual(q[b * n + i], which == 0 && i == k ? 1.0 : 0.0);
          b_[i] = Dual(qd[b * n + i], which == 1 && i == k ? 1.0 : 0.0);
          c[i] = Dual(qdd[b * n + i], which == 2 && i == k ? 1.0 : 0.0);
        }
        rnea(m, a.data(), b_.data(), c.data(), t.data());
        for (int i = 0; i < n; ++i) out[(b * n + i) * n + k] = t[i].d;
      }
    }
  }
}

// Exact d(link position)/dq, (B, 3, nv), by dual numbers through FK.
void oracle_link_pos_jac(void *h, int64_t B, const double *q, int link, double *out) {
  const Model &m = *static_cast<Model *>(h);
  int n = m.nj;
  std::vector<Dual> qq(n);
  for (int64_t b = 0; b < B; ++b)
    for (int k = 0; k < n; ++k) {
      for (int i = 0; i < n; ++i) qq[i] = Dual(q[b * n + i], i == k ? 1.0 : 0.0);
      std::vector<SE3<Dual>> oMi;
      forward_kinematics(m, qq.data(), oMi);
      SE3<Dual> M = link_pose(m, oMi, link);
      for (int r = 0; r < 3; ++r) out[(b * 3 + r) * n + k] = M.p(r, 0).d;
    }
}

// log6 and Jlog6 of a pose (p xyz | q wxyz): out6, out36 (row-major).
void oracle_log6(const double *pose7_, double *out6, double *out36) {
  SE3<double> M = se3_from_pose7(pose7_);
  V6<double> l = log6(M);
  for (int i = 0; i < 6; ++i) out6[i] = l(i, 0);
  if (out36) {
    M6<double> J = Jlog6(M);
    for (int i = 0; i < 6; ++i)
      for (int j = 0; j < 6; ++j) out36[i * 6 + j] = J(i, j);
  }
}
void oracle_log6_mat(const double *R9, const double *p3, double *out6, double *out36) {
  SE3<double> M;
  for (int i = 0; i < 3; ++i) {
    for (int j = 0; j < 3; ++j) M.R(i, j) = R9[3 * i + j];
    M.p(i, 0) = p3[i];
  }
  V6<double> l = log6(M);
  for (int i = 0; i < 6; ++i) out6[i] = l(i, 0);
  if (out36) {
    M6<double> J = Jlog6(M);
    for (int i = 0; i < 6; ++i)
      for (int j = 0; j < 6; ++j) out36[i * 6 + j] = J(i, j);
  }
}

// Dense symmetric solve x = A^-1 b by LDL^T without pivoting (A SPD, n <= 64).
static void ldlt_solve(int n, std::vector<double> A, double *b) {
  for (int j = 0; j < n; ++j) {
    for (int k = 0; k < j; ++k) A[j * n + j] -= A[j * n + k] * A[j * n + k] * A[k * n + k];
    for (int i = j + 1; i < n; ++i) {
      double s = A[i * n + j];
      for (int k = 0; k < j; ++k) s -= A[i * n + k] * A[j * n + k] * A[k * n + k];
      A[i * n + j] = s / A[j * n + j];
    }
  }
  for (int i = 0; i < n; ++i)
    for (int k = 0; k < i; ++k) b[i] -= A[i * n + k] * b[k];
  for (int i = 0; i < n; ++i) b[i] /= A[i * n + i];
  for (int i = n - 1; i >= 0; --i)
    for (int k = i + 1; k < n; ++k) b[i] -= A[k * n + i] * b[k];
}

// CLIK (pinocchio_model.cpp:228-300). q0 == nullptr -> neutral; mask == nullptr
// -> all active. Returns iterations used; q_out in user order (continuous
// joints wrapped by atan2), success flag, best error 6-vector.
int oracle_clik(void *h, int link, const double *target7, const double *q0, const int *mask,
                double eps, int max_iter, double dt, double damp, double *q_out, int *success,
                double *err_out) {
  const Model &m = *static_cast<Model *>(h);
  int n = m.nj;
  std::vector<double> q(n, 0.0), msk(n, 1.0);
  if (q0)
    for (int i = 0; i < n; ++i) q[i] = q0[i];
  if (mask)
    for (int i = 0; i < n; ++i) msk[i] = double(mask[i]);
  int jid = m.link_joint[link];
  SE3<double> l2w = se3_from_pose7(target7);
  SE3<double> l2j = se3_from<double>(m.link_place[link]);
  SE3<double> oMdes = compose(l2w, inverse(l2j));
  std::vector<double> bestQ = q;
  double minErr = 1e10;
  double bestErr[6] = {0, 0, 0, 0, 0, 0};
  int it = 0;
  bool ok = false;
  for (int i = 0;; ++i) {
    it = i;
    std::vector<SE3<double>> oMi;
    forward_kinematics(m, q.data(), oMi);
    SE3<double> oMj = jid >= 0 ? oMi[jid] : SE3<double>();
    SE3<double> iMd = compose(inverse(oMj), oMdes);
    V6<double> err = log6(iMd);
    double en = std::sqrt(dotN(err, err));
    if (en < minErr) {
      minErr = en;
      bestQ = q;
      for (int r = 0; r < 6; ++r) bestErr[r] = err(r, 0);
    }
    if (en < eps) { ok = true; break; }
    if (i >= max_iter) { ok = false; break; }
    // Jacobian of the joint in its own frame (computeJointJacobian, LOCAL).
    std::vector<V6<double>> cols;
    world_jacobian(m, oMi, cols);
    M6<double> Ainv = motionXinv(oMj);
    std::vector<double> J(6 * n, 0.0);
    M6<double> Jl = Jlog6(inverse(iMd));
    for (int k = 0; k < n; ++k) {
      if (jid < 0 || !supports(m, k, jid)) continue;
      V6<double> c = Jl * (Ainv * cols[k]);
      for (int r = 0; r < 6; ++r) J[r * n + m.vuser[k]] = -c(r, 0) * msk[m.vuser[k]];
    }
    std::vector<double> JJt(36, 0.0);
    for (int a = 0; a < 6; ++a)
      for (int b = 0; b < 6; ++b) {
        double s = 0;
        for (int k = 0; k < n; ++k) s += J[a * n + k] * J[b * n + k];
        JJt[a * 6 + b] = s + (a == b ? damp : 0.0);
      }
    double y[6];
    for (int r = 0; r < 6; ++r) y[r] = err(r, 0);
    ldlt_solve(6, JJt, y);
    for (int k = 0; k < n; ++k) {
      double v = 0;
      for (int r = 0; r < 6; ++r) v -= J[r * n + k] * y[r];
      q[k] += v * dt;  // integrate: revolute/prismatic add; continuous rotates the angle
    }
  }
  for (int j = 0; j < n; ++j) {
    int u = m.vuser[j];
    double v = bestQ[u];
    q_out[u] = m.jtype[j] == CONTINUOUS ? std::atan2(std::sin(v), std::cos(v)) : v;
  }
  *success = ok ? 1 : 0;
  for (int r = 0; r < 6; ++r) err_out[r] = bestErr[r];
  return it;
}

// Semi-implicit Euler rollout with ABA: qd += dt * aba(q, qd, tau_t); q += dt * qd.
// tau is (B, T, nv). Outputs final (q, qd).
void oracle_rollout(void *h, int64_t B, int T, double dt, const double *q0, const double *qd0,
                    const double *tau, double *qT, double *qdT) {
  const Model &m = *static_cast<Model *>(h);
  int n = m.nj;
  std::vector<double> q(n), qd(n), qdd(n);
  for (int64_t b = 0; b < B; ++b) {
    for (int i = 0; i < n; ++i) { q[i] = q0[b * n + i]; qd[i] = qd0[b * n + i]; }
    for (int t = 0; t < T; ++t) {
      aba(m, q.data(), qd.data(), tau + (b * T + t) * n, qdd.data());
      for (int i = 0; i < n; ++i) {
        qd[i] += dt * qdd[i];
        q[i] += dt * qd[i];
      }
    }
    for (int i = 0; i < n; ++i) { qT[b * n + i] = q[i]; qdT[b * n + i] = qd[i]; }
  }
}

// Directional derivative of the rollout's final state: given seeds (dq0, dqd0,
// dtau) returns (d qT, d qdT) exactly (forward mode). Used to check adjoints.
void oracle_rollout_jvp(void *h, int T, double dt, const double *q0, const double *qd0,
                        const double *tau, const double *dq0, const double *dqd0,
                        const double *dtau, double *dqT, double *dqdT) {
  const Model &m = *static_cast<Model *>(h);
  int n = m.nj;
  std::vector<Dual> q(n), qd(n), qdd(n), tt(n);
  for (int i = 0; i < n; ++i) { q[i] = Dual(q0[i], dq0[i]); qd[i] = Dual(qd0[i], dqd0[i]); }
  for (int t = 0; t < T; ++t) {
    for (int i = 0; i < n; ++i) tt[i] = Dual(tau[t * n + i], dtau[t * n + i]);
    aba(m, q.data(), qd.data(), tt.data(), qdd.data());
    for (int i = 0; i < n; ++i) {
      qd[i] += Dual(dt) * qdd[i];
      q[i] += Dual(dt) * qd[i];
    }
  }
  for (int i = 0; i < n; ++i) { dqT[i] = q[i].d; dqdT[i] = qd[i].d; }
}

// CPU baseline workload: FK (all link poses) + full WORLD Jacobian + RNEA per
// sample, FP64, statically partitioned over `threads`. Returns seconds.
double oracle_fk_jac_rnea_timed(void *h, int64_t B, const double *q, const double *qd,
                                const double *qdd, double *poses, double *J, double *tau,
                                int threads) {
  const Model &m = *static_cast<Model *>(h);
  int n = m.nj, L = m.nlinks;
  auto t0 = std::chrono::steady_clock::now();
  parallel_for(B, threads, [&](int64_t b) {
    std::vector<SE3<double>> oMi;
    std::vector<V6<double>> cols;
    forward_kinematics(m, q + b * n, oMi);
    for (int l = 0; l < L; ++l) pose7(link_pose(m, oMi, l), poses + (b * L + l) * 7);
    world_jacobian(m, oMi, cols);
    for (int j = 0; j < n; ++j)
      for (int r = 0; r < 6; ++r) J[(b * 6 + r) * n + m.vuser[j]] = cols[j](r, 0);
    rnea(m, q + b * n, qd + b * n, qdd + b * n, tau + b * n);
  });
  auto t1 = std::chrono::steady_clock::now();
  return std::chrono::duration<double>(t1 - t0).count();
}

// Exact operation counts of one FK + all link poses + full Jacobian + RNEA
// evaluation in this textbook formulation: out = {adds, muls, divs, sqrts, trig}.
void oracle_opcount_fk_jac_rnea(void *h, const double *q, const double *qd, const double *qdd,
                                uint64_t *out) {
  const Model &m = *static_cast<Model *>(h);
  int n = m.nj, L = m.nlinks;
  std::vector<OpCount> a(n), b(n), c(n), t(n);
  for (int i = 0; i < n; ++i) { a[i] = q[i]; b[i] = qd[i]; c[i] = qdd[i]; }
  OpCount::reset();
  std::vector<SE3<OpCount>> oMi;
  std::vector<V6<OpCount>> cols;
  forward_kinematics(m, a.data(), oMi);
  OpCount pz[7];
  for (int l = 0; l < L; ++l) pose7(link_pose(m, oMi, l), pz);
  world_jacobian(m, oMi, cols);
  rnea(m, a.data(), b.data(), c.data(), t.data());
  out[0] = OpCount::adds; out[1] = OpCount::muls; out[2] = OpCount::divs;
  out[3] = OpCount::sqrts; out[4] = OpCount::trig;
}

}  // extern "C"
