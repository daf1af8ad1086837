// Kernel launchers (host-callable), one per algorithm, templated on the
// arithmetic type. All pointers are device pointers in the user dof order:
// q/qd/qdd/tau (B, nv); poses (B, L, 7) = p xyz | q wxyz; J (B, 6, nv);
// per-link Jacobians (B, L, 6, nv); M/C (B, nv, nv).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "tables.h"

namespace sapamd {

enum KinFlags : int {
  KF_POSES = 1, KF_JFULL = 2, KF_JLINKS = 4, KF_RNEA = 8, KF_CRBA = 64,
  // RNEA variants (passive forces): no gravity term, qd = 0, qdd = 0 (the
  // corresponding input pointer is then not read)
  KF_NOGRAV = 256, KF_QD0 = 512, KF_QDD0 = 1024
};

// Outputs and the layout of the batched arrays. Strides are in elements; 0
// means contiguous: q/qd/qdd/tau rows of nv, poses (B, L, 7). Non-zero strides
// let a caller pass padded buffers without copies (a simulator's
// [n_art, max_dof] joint state, a [n_art, max_links, 13] body-state buffer whose
// first 7 floats per link are p xyz | q wxyz).
template <class T> struct KinOut {
  T *poses = nullptr, *jfull = nullptr, *jw_links = nullptr, *jl_links = nullptr, *tau = nullptr;
  T *M = nullptr;  // (B, nv, nv), tree kernels only
  int64_t ld_q = 0, ld_qd = 0, ld_qdd = 0, ld_tau = 0;
  int64_t pose_env = 0, pose_link = 0;
};

// One tree type of a heterogeneous batch: its table and buffers (device
// pointers, user order), its lanes per environment and its first block.
struct HeteroSegment {
  const ModelTable<float> *tab;
  int32_t G, nj;
  int64_t B, block0;
  const float *q, *qd, *qdd;
  float *poses, *J, *tau, *M;
};
// Lanes-per-joint kernels for trees of <= 32 joints (FP32): FK + Jacobian +
// RNEA + CRBA in one pass (flags select outputs).
int tree_group(int nj);
// G = lanes per environment (8, 16 or 32, >= nj; 0 = tree_group(nj)).
// ev0 / ev1 (optional): events that receive the kernel's own start and stop
// (hipExtLaunchKernel), for kernel timing without a profiler.
hipError_t launch_tree_fused(const ModelTable<float> *tab, int nj, int64_t B, int flags, const float *q,
                             const float *qd, const float *qdd, const KinOut<float> &out, hipStream_t s, int G = 0,
                             hipEvent_t ev0 = nullptr, hipEvent_t ev1 = nullptr);
template <class T> struct VjpArgs;
// Tree version of the fused forward + reverse pass (pose gradients need one
// user link per joint).
// Link poses (and, with qd, centre-of-mass and angular velocities) of every
// user link into body-state rows of 13 floats (p | q wxyz | v | w) every
// link_stride floats within rows of env_stride floats; row 0 of each
// environment (the root link) is read first as the articulation's world pose
// and velocity. Trees of 1 to 32 joints; q / qd rows every ld_q / ld_qd floats.
hipError_t launch_link_state(const ModelTable<float> *tab, int nj, int64_t B, const float *q, int64_t ld_q,
                             const float *qd, int64_t ld_qd, float *body, int64_t env_stride, int64_t link_stride,
                             hipStream_t s);
hipError_t launch_tree_vjp(const ModelTable<float> *tab, int nj, int64_t B, const float *q, const float *qd,
                           const float *qdd, VjpArgs<float> a, hipStream_t s);
// All segments in one launch (passed by value, up to kMaxSegments); the grid is
// the sum of the segments' block counts (256-thread blocks).
constexpr int kMaxSegments = 8;
struct HeteroArgs {
  HeteroSegment seg[kMaxSegments];
  int32_t nseg;
};
hipError_t launch_hetero(const HeteroArgs &args, int64_t total_blocks, int flags, hipStream_t s);

constexpr int kMaxGenericJoints = 64;

template <class T>
hipError_t launch_kinematics(const ModelTable<T> *tab, int nj, int64_t B, int flags, const T *q, const T *qd,
                             const T *qdd, KinOut<T> out, hipStream_t s);
template <class T>
hipError_t launch_crba(const ModelTable<T> *tab, int nj, int64_t B, const T *q, T *M, hipStream_t s);
template <class T>
hipError_t launch_coriolis(const ModelTable<T> *tab, int nj, int64_t B, const T *q, const T *qd, T *C,
                           hipStream_t s);
template <class T>
hipError_t launch_aba(const ModelTable<T> *tab, int nj, int64_t B, const T *q, const T *qd, const T *tau, T *qdd,
                      hipStream_t s);
// Local (body-frame) Jacobian of one link: Ad(link2world)^-1 J (B, 6, nv).
template <class T>
hipError_t launch_link_local_jacobian(const ModelTable<T> *tab, int nj, int64_t B, int link, const T *q, T *J,
                                      hipStream_t s);

struct ClikParams {
  int link;            // user link index
  const int32_t *chain;  // internal joint ids root -> link joint (device), length nchain
  int nchain;
  double eps, dt, damp;
  int max_iter;
  const int32_t *mask;  // (nv) or (B, nv) int, nullptr = all active
  int64_t mask_ld;      // 0 = shared mask, nv = per problem
  int32_t *counter;     // work-queue counter, zeroed before the launch
  int persistent_waves; // 0 = one lane per problem
  // Multi-start: problem p solves target p % ntargets from start p / ntargets
  // (start-major order); solved[t] holds the lowest start index that converged
  // for target t (INT32_MAX before), and a problem whose target already
  // converged from a lower start is skipped (success 0, iterations -1).
  int64_t ntargets = 0;  // 0 = one target per problem
  int32_t *solved = nullptr;
};
// CLIK (pinocchio_model.cpp:228-300) for B problems sharing one target link.
// target (B, 7) p | wxyz; q0 (B, nv) or nullptr (neutral); outputs q (B, nv),
// success (B), err (B, 6), iters (B).
template <class T>
hipError_t launch_clik(const ModelTable<T> *tab, int nj, int64_t B, const ClikParams &P, const T *target,
                       const T *q0, T *q_out, int32_t *success, T *err, int32_t *iters, hipStream_t s);

// CLIK problems of several models in one persistent launch (FP32). Segment s
// owns global problems [first, first + B); A.counter is zeroed before the launch.
struct ClikSegment {
  const ModelTable<float> *tab;
  ClikParams P;
  const float *target, *q0;
  float *q_out, *err;
  int32_t *success, *iters;
  int64_t B, first;
};
// State of a problem handed from the first CLIK phase to the second: the
// iterate about to be evaluated, the iteration count and the best iterate so
// far, per chain joint.
struct ClikCont {
  int64_t prob;  // problem index over the launch's segments
  int32_t it;
  float minErr;
  float bestErr[6];
  float q[16], qbest[16];
};
struct HeteroClikArgs {
  ClikSegment seg[kMaxSegments];
  int32_t nseg;
  int64_t total;
  int32_t *counter;
  // Two-phase lanes-per-joint CLIK: problems still running after `cap`
  // iterations go to `cont` and finish in a second launch that takes them
  // from cont_queue. Those still far from their target (likely to run to the
  // iteration limit) fill `cont` from the front (count in cont_count), the
  // others from the back (count in cont_easy), so the second launch starts
  // the long ones first. cap = 0: one phase.
  ClikCont *cont = nullptr;
  int32_t *cont_count = nullptr, *cont_queue = nullptr, *cont_easy = nullptr;
  int32_t cap = 0;
};
hipError_t launch_clik_hetero(const HeteroClikArgs &A, int max_chain, int waves, hipStream_t s);
// Lanes-per-joint CLIK (FP32, chains of <= 16 joints) in two phases:
// persistent blocks of 256 threads pulling problems from A.counter, first one
// problem per 4 lanes x 4 chain joints for at most kClikPhase1Iters
// iterations, then the problems still running one per 8 lanes x 2 joints
// (shorter iterations for the long tail). A.counter, A.cont_count,
// A.cont_queue and A.cont_easy are zeroed before the launch; A.cont holds
// A.total records.
// Serves the per-model batched and the mixed-tree launches;
// clik_lanes_group returns the first phase's group size, 0 above 16 chain
// joints.
// The schedule: phase one runs iterations [0, cap) at g1 lanes x (16 / g1)
// chain joints per problem, phase two the problems still running at g2 lanes
// (g1, g2 in {4, 8, 16}); cap = 0 runs every problem to its end in phase one.
constexpr int kClikPhase1Iters = 128;
struct ClikSchedule {
  int cap = kClikPhase1Iters, g1 = 4, g2 = 8;
};
bool clik_schedule_valid(const ClikSchedule &s);
int clik_lanes_group(int max_chain);
hipError_t launch_clik_lanes(const HeteroClikArgs &A, int max_chain, int cus, const ClikSchedule &sch,
                             hipStream_t s);

// Lanes-per-joint fused FK + Jacobian + RNEA for serial chains of <= 16 joints
// (FP32), G lanes per environment x C joints per lane. Returns
// hipErrorNotSupported for other trees or unsupported shapes.
int fast_chain_group(int nj);
int chain_shape_count();
void chain_shape(int i, int *G, int *C);
void default_chain_shape(int nj, int *G, int *C);
// ev0 / ev1 as for launch_tree_fused.
hipError_t launch_chain_fused(const ModelTable<float> *tab, int nj, int64_t B, int flags, const float *q,
                              const float *qd, const float *qdd, KinOut<float> out, hipStream_t s, int G, int C,
                              hipEvent_t ev0 = nullptr, hipEvent_t ev1 = nullptr);

// Reverse mode (vector-Jacobian products) of FK link poses and RNEA.
// VJP mode: upstream gradients gtau (B, nv) and/or gposes (B, L, 7).
// Loss mode (tau_star / p_target given): the kernel forms them itself for
//   loss = sum_i (tau_i - tau*_i)^2 + |p_ee - t|^2,
// writes the per-environment loss and returns its gradients.
// Gradient outputs: dq always; dqd, dqdd when the RNEA term is present.
template <class T> struct VjpArgs {
  const T *gtau = nullptr;      // (B, nv)
  const T *gposes = nullptr;    // (B, L, 7)
  const T *tau_star = nullptr;  // (B, nv), loss mode
  const T *p_target = nullptr;  // (B, 3), loss mode
  int ee_link = -1;             // user link of p_target
  T *loss = nullptr;            // (B), loss mode
  T *tau = nullptr;             // (B, nv), optional forward output
  T *dq = nullptr, *dqd = nullptr, *dqdd = nullptr;
};
template <class T>
hipError_t launch_kin_vjp(const ModelTable<T> *tab, int nj, int64_t B, const T *q, const T *qd, const T *qdd,
                          VjpArgs<T> a, hipStream_t s);
// Lanes-per-joint version of the same for serial chains (FP32).
hipError_t launch_chain_vjp(const ModelTable<float> *tab, int nj, int64_t B, const float *q, const float *qd,
                            const float *qdd, VjpArgs<float> a, hipStream_t s, int G, int C);

// One semi-implicit Euler step with ABA, in place on q, qd (B, nv); tau_t rows
// at tau_stride; when save != nullptr the pre-step state (q | qd) is written
// to save rows at save_stride. tab64 != nullptr selects FP64 factorized
// forward dynamics (CRBA + RNEA bias + LTDL) instead of FP32 ABA.
template <class T>
hipError_t launch_rollout_step(const ModelTable<T> *tab, const ModelTable<double> *tab64, int nj, int64_t B, T dt,
                               T *q, T *qd, const T *tau, int64_t tau_stride, T *save, int64_t save_stride,
                               hipStream_t s);
// Adjoint of one step: state rows (q | qd) before the step, aq/aqd (B, nv)
// hold the adjoint of the post-step state on entry and of the pre-step state
// on exit; gtau_t rows (at tau_stride) receive dL/dtau_t.
template <class T>
hipError_t launch_rollout_step_vjp(const ModelTable<T> *tab, const ModelTable<double> *tab64, int nj, int64_t B,
                                   T dt, const T *state, int64_t state_stride, const T *tau, int64_t tau_stride,
                                   T *aq, T *aqd, T *gtau, hipStream_t s);
// Batched forward dynamics in FP64 (CRBA + RNEA bias + LTDL), FP32 in and out.
hipError_t launch_fd_ldlt(const ModelTable<double> *tab64, int nj, int64_t B, const float *q, const float *qd,
                          const float *tau, float *qdd, hipStream_t s);
// Lanes-per-joint forward dynamics for trees of <= 32 joints: CRBA + RNEA
// bias + LDL^T in LDS, in FP64 when tab64 is given, else FP32 (FP32 inputs,
// outputs and state either way). mode 0: out = qdd (B, nv); mode 1: one rollout
// step in place on q, qd (save as launch_rollout_step). hipErrorNotSupported
// above 32 joints.
// rows: the factorization size, nj, or the largest root subtree when the
// factorization runs per root subtree (dyn_factor_plan); 0 = nj.
int tree_dyn_group(int nj);
// The compile-time elimination bound for a factorization of `rows` rows in a
// group of G lanes: the smallest of 8, 12, 16, 20, 24, 32 >= rows (<= G).
int tree_dyn_steps(int rows, int G);
// Whether the factorization runs per root subtree (M is block diagonal over
// them) and its size: per subtree when there are several, it shortens the
// elimination, and every block's rows stay in the group (first lane + steps <= G).
struct DynFactorPlan {
  bool blocks;
  int rows;
};
DynFactorPlan dyn_factor_plan(const int *parent, int nj);
hipError_t launch_tree_step(const ModelTable<float> *tab, const ModelTable<double> *tab64, int nj, int64_t B,
                            int mode, float dt, float *q, float *qd, const float *tau, int64_t tau_stride, float *out,
                            float *save, int64_t save_stride, hipStream_t s, int rows = 0);
hipError_t launch_tree_step_vjp(const ModelTable<float> *tab, const ModelTable<double> *tab64, int nj, int64_t B,
                                float dt, const float *state, int64_t state_stride, const float *tau,
                                int64_t tau_stride, float *aq, float *aqd, float *gtau, hipStream_t s, int rows = 0);
// T rollout steps in one launch (state in registers; ckpt (B, ceil(T/K), 2 nv)
// optional), and its checkpointed adjoint in one launch (seg (B, K, 2 nv)).
hipError_t launch_tree_rollout(const ModelTable<float> *tab, const ModelTable<double> *tab64, int nj, int64_t B,
                               int T, int K, float dt, float *q, float *qd, const float *tau, float *ckpt,
                               hipStream_t s, int rows = 0);
hipError_t launch_tree_rollout_vjp(const ModelTable<float> *tab, const ModelTable<double> *tab64, int nj, int64_t B,
                                   int T, int K, float dt, const float *tau, const float *ckpt, float *seg, float *aq,
                                   float *aqd, float *gtau, hipStream_t s, int rows = 0);
// One model of a mixed-tree rollout launch: its table (ModelTable<float>, or
// ModelTable<double> when fp64), joints and factorization rows (as for
// launch_tree_rollout), batch, first block and buffers as in
// launch_tree_rollout (q, qd, tau, ckpt) / launch_tree_rollout_vjp (tau, ckpt,
// seg, aq, aqd, gtau). variant and block0 are filled by launch_rollout_mixed.
struct RolloutSegment {
  const void *tab;
  int32_t nj, rows, fp64, variant;
  int64_t B, block0;
  float *q, *qd;
  const float *tau, *ckpt;
  float *ckpt_out, *seg, *aq, *aqd, *gtau;
};
struct RolloutMixArgs {
  RolloutSegment seg[kMaxSegments];
  int32_t nseg, T, K;
  float dt;
};
// Rollouts (vjp = false: forward with checkpoints into ckpt_out) or their
// adjoints (vjp = true) of up to kMaxSegments models in ONE launch; the
// segments are ordered heaviest first so the longest-running blocks start
// first. Same bits per environment as the per-model launches.
hipError_t launch_rollout_mixed(RolloutMixArgs A, bool vjp, hipStream_t s);
// rows of (q | qd) at stride -> q, qd (B, nv)
hipError_t launch_restore_state(int64_t B, int nv, const float *src, int64_t stride, float *q, float *qd,
                                hipStream_t s);

// dst[0:n] = value (a kernel, not hipMemsetAsync, so it is safe under graph capture)
hipError_t launch_fill_i32(int32_t *dst, int32_t value, int64_t n, hipStream_t s);
// B random configurations (B, nv), entry (b, u) uniform in [lo[u], hi[u]) from
// Philox4x32-10 at counter (b, u) under the 64-bit seed.
constexpr int kRandomMaxDofs = 64;
struct RandomLimits {
  double lo[kRandomMaxDofs], hi[kRandomMaxDofs];
};
template <class T>
hipError_t launch_random_qpos(int64_t B, int nv, uint64_t seed, const RandomLimits &lim, T *out, hipStream_t s);

}  // namespace sapamd
