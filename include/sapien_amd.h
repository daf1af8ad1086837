/* sapien_amd C ABI — MI355X (gfx950) articulated-body kinematics and dynamics.
 *
 * Drop-in boundary for SAPIEN's PinocchioModel path. Every entry point
 * returns an int status (SAPAMD_OK = 0); on failure the message is available
 * from sapamd_last_error() (thread-local). No C++ exception crosses the ABI,
 * no torch type appears in a signature, and the library never frees a buffer
 * it did not allocate.
 *
 * Two families:
 *   - per-call entry points with the reference's method semantics
 *     (host arrays, FP64, one configuration per call, results cached in the
 *     model exactly where the reference caches them in pinocchio::Data);
 *   - batched entry points over caller-owned device buffers (FP32, row-major,
 *     contiguous, user dof order), enqueued on a caller-supplied hipStream_t
 *     passed as void* (nullptr = the null stream).
 *
 * Layouts (batched): q/qd/qdd/tau (B, nv); link poses (B, L, 7) as
 * p xyz | quaternion wxyz; full Jacobian (B, 6, nv) with rows
 * [linear; angular] in the WORLD frame (linear rows = velocity of the point at
 * the world origin); mass/Coriolis matrices (B, nv, nv).
 *
 * Reference files are cited relative to the SAPIEN repository root.
 */
#ifndef SAPIEN_AMD_H
#define SAPIEN_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  SAPAMD_OK = 0,
  SAPAMD_INVALID_ARGUMENT = 1, /* Python maps to ValueError   (pinocchio_model.cpp:105,122,136) */
  SAPAMD_RUNTIME_ERROR = 2,    /* Python maps to RuntimeError (pinocchio_model.cpp:33-36,64-66,70,113,151) */
  SAPAMD_DEVICE_ERROR = 3      /* HIP runtime failure */
};

typedef struct sapamd_model sapamd_model;

typedef struct {
  int32_t nv;          /* user dofs (model.nv) */
  int32_t nq;          /* configuration size with continuous joints as (cos, sin) (model.nq) */
  int32_t njoints;     /* movable joints, universe excluded */
  int32_t nlinks;      /* links given to set_link_order */
  int32_t max_depth;   /* longest root-to-leaf joint chain */
  int32_t chain;       /* 1 when the joints form one serial chain */
  int32_t supported;   /* 0 when the URDF holds floating/planar joints */
  int32_t order_set;   /* 1 after a successful set_joint_order */
  int32_t device;      /* HIP device holding the model tables */
} sapamd_info;

/* Thread-local message of the last failed call on this thread. */
const char *sapamd_last_error(void);
/* Library version string. */
const char *sapamd_version(void);

/* ---------------------------------------------------------------- model */
/* PinocchioModel(urdf, gravity): pinocchio_model.cpp:39-47 (fromURDFXML), :310-312.
 * Builds the model on the current HIP device. */
int sapamd_model_create(const char *urdf, const double gravity[3], sapamd_model **out);
void sapamd_model_destroy(sapamd_model *m);
int sapamd_model_info(const sapamd_model *m, sapamd_info *out);
/* set_joint_order: pinocchio_model.cpp:99-128. */
int sapamd_set_joint_order(sapamd_model *m, const char *const *names, int32_t n);
/* set_link_order: pinocchio_model.cpp:130-140. */
int sapamd_set_link_order(sapamd_model *m, const char *const *names, int32_t n);
/* Joint limits in the user order (lower, upper); continuous joints give (-pi, pi). */
int sapamd_joint_limits(const sapamd_model *m, double *lower, double *upper);
/* Per user dof: 1 when the joint is continuous (returned q is atan2-wrapped). */
int sapamd_continuous_mask(const sapamd_model *m, int32_t *out);

/* ---------------------------------------------------------------- per-call (reference semantics, FP64) */
/* compute_forward_kinematics: pinocchio_model.cpp:146-148 (caches every link pose). */
int sapamd_compute_forward_kinematics(sapamd_model *m, const double *qpos, int64_t n);
/* get_link_pose: pinocchio_model.cpp:150-161 -> p xyz | q wxyz, FP32 like sapien::Pose. */
int sapamd_get_link_pose(const sapamd_model *m, int32_t link, float pose7[7]);
/* compute_full_jacobian: pinocchio_model.cpp:163-165. */
int sapamd_compute_full_jacobian(sapamd_model *m, const double *qpos, int64_t n);
/* get_link_jacobian: pinocchio_model.cpp:167-186; local != 0 gives the body (link)
 * frame Jacobian, as the reference code does (its docstring at :355-362 is inverted).
 * out: 6 x nv row-major. */
int sapamd_get_link_jacobian(const sapamd_model *m, int32_t link, int32_t local, double *out);
/* compute_single_link_local_jacobian: pinocchio_model.cpp:188-199. out 6 x nv. */
int sapamd_compute_single_link_local_jacobian(sapamd_model *m, const double *qpos, int64_t n, int32_t link,
                                              double *out);
/* compute_generalized_mass_matrix: pinocchio_model.cpp:201-206 (full symmetric). */
int sapamd_compute_generalized_mass_matrix(sapamd_model *m, const double *qpos, int64_t n, double *out);
/* compute_coriolis_matrix: pinocchio_model.cpp:208-212 (Christoffel-consistent C). */
int sapamd_compute_coriolis_matrix(sapamd_model *m, const double *qpos, const double *qvel, int64_t n,
                                   double *out);
/* compute_inverse_dynamics: pinocchio_model.cpp:214-219 (RNEA). */
int sapamd_compute_inverse_dynamics(sapamd_model *m, const double *qpos, const double *qvel, const double *qacc,
                                    int64_t n, double *out);
/* compute_forward_dynamics: pinocchio_model.cpp:221-226 (ABA). */
int sapamd_compute_forward_dynamics(sapamd_model *m, const double *qpos, const double *qvel, const double *qf,
                                    int64_t n, double *out);
/* get_random_qpos: pinocchio_model.cpp:142-144 (getRandomConfiguration, user
 * order): out (nv) uniform within the joint limits (continuous joints in
 * [-pi, pi)), drawn on the GPU by the generator of sapamd_random_qpos_batched
 * (row 0 of the batch for this seed). */
int sapamd_get_random_qpos(sapamd_model *m, uint64_t seed, double *out);
/* compute_inverse_kinematics: pinocchio_model.cpp:228-300. pose7 = p xyz | q wxyz;
 * initial_qpos may be NULL (n0 = 0: neutral); active_qmask may be NULL (nm = 0:
 * all active). Outputs q (nv), success, best error (6). */
int sapamd_compute_inverse_kinematics(sapamd_model *m, int32_t link, const float pose7[7],
                                      const double *initial_qpos, int64_t n0, const int32_t *active_qmask,
                                      int64_t nm, double eps, int32_t max_iterations, double dt, double damp,
                                      double *q_out, int32_t *success, double err6[6]);

/* ---------------------------------------------------------------- batched device path (FP32) */
/* Kernel family used by the batched kinematics entry points (no reference
 * counterpart; the reference has no GPU path). 0 = auto (lanes-per-joint for
 * serial chains of <= 16 joints and for trees of <= 32 joints, else
 * thread-per-environment), 1 = thread-per-environment, 2 = lanes-per-joint
 * (fails for larger trees).
 * Not synchronized with launches in flight on other threads. */
enum { SAPAMD_KERNEL_AUTO = 0, SAPAMD_KERNEL_THREAD_PER_ENV = 1, SAPAMD_KERNEL_LANES_PER_JOINT = 2 };
int sapamd_set_kernel_variant(sapamd_model *m, int32_t variant);
/* The variant the batched kinematics entry points would launch now (1 or 2). */
int sapamd_get_kernel_variant(const sapamd_model *m, int32_t *variant);
/* Shapes of the lanes-per-joint chain kernel: G lanes per environment x C joints
 * per lane. Writes up to cap (G, C) pairs and returns how many exist. */
int sapamd_chain_shapes(int32_t *G, int32_t *C, int32_t cap);
/* Use shape (G, C) for batches of at most max_batch environments (the
 * autotuner's result; the smallest covering entry wins, larger batches use the
 * last entry's default). G = C = 0 clears every entry. */
int sapamd_set_chain_shape(sapamd_model *m, int64_t max_batch, int32_t G, int32_t C);
/* The shape a batch of B environments would launch with. */
int sapamd_get_chain_shape(const sapamd_model *m, int64_t B, int32_t *G, int32_t *C);

/* Lanes per environment of the lanes-per-joint tree kernels (batched
 * kinematics of trees that are not short chains, and this model's segment of
 * a heterogeneous launch): 8, 16 or 32, at least the joint count; 0 = the
 * smallest that holds the joints (the autotuner's setting; no reference
 * counterpart). */
int sapamd_set_tree_group(sapamd_model *m, int32_t G);
int sapamd_get_tree_group(const sapamd_model *m, int32_t *G);
/* Schedule of the lanes-per-joint CLIK launches, library-wide so that a
 * problem gives the same bits in a per-model and a mixed-tree launch: the
 * first launch runs iterations [0, cap) at g1 lanes x (16 / g1) chain joints
 * per problem, the second the problems still running at g2 lanes; cap = 0 runs
 * every problem to its end in the first (g1, g2 in {4, 8, 16}; default 128, 4,
 * 8). The autotuner's setting; no reference counterpart. */
int sapamd_set_clik_schedule(int32_t cap, int32_t g1, int32_t g2);
int sapamd_get_clik_schedule(int32_t *cap, int32_t *g1, int32_t *g2);

/* Fused FK (every link pose) + full WORLD Jacobian + RNEA over B environments;
 * the throughput-critical entry point. Any of poses/J/tau may be NULL to skip it. */
int sapamd_fk_jac_rnea_batched(const sapamd_model *m, int64_t B, const float *q, const float *qd,
                               const float *qdd, float *poses, float *J, float *tau, void *stream);
/* The same launch with the kernel's own start and stop recorded into two
 * caller-created hipEvent_t (hipExtLaunchKernel), so hipEventElapsedTime gives
 * the kernel's execution time without the dispatch gaps around it: kernel
 * timing without a profiler (lanes-per-joint kernels only; not for graph
 * capture). No reference counterpart. */
int sapamd_fk_jac_rnea_batched_timed(const sapamd_model *m, int64_t B, const float *q, const float *qd,
                                     const float *qdd, float *poses, float *J, float *tau, void *start_event,
                                     void *stop_event, void *stream);
/* The same over padded buffers, without copies: q/qd/qdd/tau rows every *_ld
 * floats (0 = nv), link poses every pose_link_stride floats (0 = 7) within
 * rows of pose_env_stride floats (0 = L * 7); J stays (B, 6, nv). This reads a
 * GPU simulator's joint state [n_art, max_dof] in place
 * (src/physx/physx_system.cpp:1076-1101). Poses are in the model's root frame,
 * as get_link_pose returns them; sapamd_link_state_strided writes world-frame
 * link poses and velocities into a simulator's body-state buffer. */
int sapamd_fk_jac_rnea_strided(const sapamd_model *m, int64_t B, const float *q, int64_t q_ld, const float *qd,
                               int64_t qd_ld, const float *qdd, int64_t qdd_ld, float *poses,
                               int64_t pose_env_stride, int64_t pose_link_stride, float *J, float *tau,
                               int64_t tau_ld, void *stream);
/* Link poses and velocities of B articulations written into a GPU
 * simulator's body-state buffer, the work of PhysxSystemGpu::
 * gpuUpdateArticulationKinematics followed by gpuFetchArticulationLinkPose /
 * gpuFetchArticulationLinkVel (src/physx/physx_system.cpp:585-618,711-721):
 * body rows of 13 floats p xyz | q wxyz | v xyz | w xyz (SapienBodyData,
 * src/physx/physx_system.cuh:89-94), link l of articulation b at
 * body + b * body_env_stride + l * body_link_stride (0 = L * 13 and 13; the
 * [n_art, max_links, 13] layout of src/physx/physx_system.cpp:1048-1051).
 * Row 0 of each articulation holds user link 0 (the simulator's root link)
 * and is read first as its world pose (world minus the scene offset,
 * src/physx/physx_system.cu:85-99) and, with qd, its velocity, which place the
 * articulation; then every user link's row receives its world pose and, with
 * qd, the linear velocity of its centre of mass and its angular velocity (world
 * frame), row 0 included (rewritten to the same values up to rounding). q / qd rows every q_ld / qd_ld floats (0 = nv); qd may be NULL
 * (poses only). Trees of 1 to 32 movable joints. */
int sapamd_link_state_strided(const sapamd_model *m, int64_t B, const float *q, int64_t q_ld, const float *qd,
                              int64_t qd_ld, float *body, int64_t body_env_stride, int64_t body_link_stride,
                              void *stream);
/* Batched get_random_qpos (pinocchio_model.cpp:142-144,
 * python/py_package/wrapper/pinocchio_model.py:113-114): out (B, nv), entry
 * (b, u) uniform within the joint limits of user dof u (continuous joints in
 * [-pi, pi)), from Philox4x32-10 at counter (b, u, 0, 0) under the key
 * (seed low, seed high 32 bits): the same seed gives the same batch. */
int sapamd_random_qpos_batched(const sapamd_model *m, int64_t B, uint64_t seed, float *out, void *stream);
/* Passive joint forces, as PhysxArticulation::computePassiveForce
 * (src/physx/articulation.cpp:400-418): gravity compensation g(q) and/or
 * Coriolis and centrifugal forces C(q, qd) qd, i.e. RNEA(q, qd, 0) with the
 * chosen terms (gravity = 0 drops g; coriolis_and_centrifugal = 0 reads no qd). */
int sapamd_passive_force_batched(const sapamd_model *m, int64_t B, const float *q, int64_t q_ld, const float *qd,
                                 int64_t qd_ld, int32_t gravity, int32_t coriolis_and_centrifugal, float *tau,
                                 int64_t tau_ld, void *stream);
/* Batched compute_forward_kinematics + get_link_pose for all links: poses (B, L, 7). */
int sapamd_fk_batched(const sapamd_model *m, int64_t B, const float *q, float *poses, void *stream);
/* Batched compute_full_jacobian: (B, 6, nv). */
int sapamd_jacobian_batched(const sapamd_model *m, int64_t B, const float *q, float *J, void *stream);
/* Batched get_link_jacobian for every link (world and body frame): (B, L, 6, nv) each; either may be NULL. */
int sapamd_link_jacobians_batched(const sapamd_model *m, int64_t B, const float *q, float *J_world, float *J_local,
                                  void *stream);
/* Batched compute_inverse_dynamics: tau (B, nv). */
int sapamd_rnea_batched(const sapamd_model *m, int64_t B, const float *q, const float *qd, const float *qdd,
                        float *tau, void *stream);
/* Batched compute_generalized_mass_matrix: (B, nv, nv). */
int sapamd_crba_batched(const sapamd_model *m, int64_t B, const float *q, float *M, void *stream);
/* Batched compute_coriolis_matrix: (B, nv, nv). */
int sapamd_coriolis_batched(const sapamd_model *m, int64_t B, const float *q, const float *qd, float *C,
                            void *stream);
/* Batched compute_forward_dynamics (ABA): qdd (B, nv). */
int sapamd_aba_batched(const sapamd_model *m, int64_t B, const float *q, const float *qd, const float *tau,
                       float *qdd, void *stream);
/* Batched compute_inverse_kinematics for B problems on one link: target (B, 7),
 * q0 (B, nv) or NULL (neutral), mask (nv) or NULL; outputs q (B, nv),
 * success (B), err (B, 6), iterations (B, may be NULL). Problems exit
 * independently; a persistent launch refills finished lanes from a work queue. */
int sapamd_clik_batched(const sapamd_model *m, int64_t B, int32_t link, const float *target, const float *q0,
                        const int32_t *mask, double eps, int32_t max_iterations, double dt, double damp,
                        float *q_out, int32_t *success, float *err, int32_t *iterations, void *stream);

/* Allocation-free sapamd_clik_batched: the same arguments, kernels and results,
 * but the queue counters and hand-over records come from a fixed scratch array
 * in the library's device memory (about 11 MB per GPU, reserved when the
 * library loads and held until the process ends) instead of memory the model
 * allocates: the call makes no hipMalloc / hipFree and allocates nothing
 * itself. No counterpart in the reference: for callers that cannot allocate (code under
 * HIP graph capture, tight serving loops). At most 65536 problems per call
 * (more fail with SAPAMD_INVALID_ARGUMENT). Every call on one GPU shares that
 * GPU's scratch, so calls that overlap on one GPU (other streams, other
 * threads) race; the counters are cleared by a fill kernel at each call. */
int sapamd_clik_batched_noalloc(const sapamd_model *m, int64_t B, int32_t link, const float *target,
                                const float *q0, const int32_t *mask, double eps, int32_t max_iterations, double dt,
                                double damp, float *q_out, int32_t *success, float *err, int32_t *iterations,
                                void *stream);

/* Heterogeneous batch in ONE launch: models[i] (trees of 1 to 32 joints, at
 * most 8 models) each with B[i] environments and its own device buffers; per
 * model any of poses/J/tau/M may be NULL (arrays of pointers may be NULL to
 * skip that output everywhere). Computes FK link poses, WORLD Jacobian, RNEA
 * and CRBA (no reference counterpart: the reference evaluates one model per call). */
int sapamd_hetero_batched(const sapamd_model *const *models, int32_t nmodels, const int64_t *B,
                          const float *const *q, const float *const *qd, const float *const *qdd,
                          float *const *poses, float *const *J, float *const *tau, float *const *M, void *stream);

/* Multi-start CLIK (the mplib-style Planner.IK(target, init_qpos, n_init_qpos)
 * use of compute_inverse_kinematics, docs/source/tutorial/motion_planning/
 * inverse_kinematics.rst:7-19): B targets (B, 7), nstart starts each; q0 is
 * (nstart * B, nv) start-major (start k of target t at row k * B + t), the
 * outputs have the same layout. A start is skipped (success 0, iterations -1,
 * error 1e30) once a lower start of its target converged; solved (B) receives
 * the lowest converged start per target (INT32_MAX when none did). */
int sapamd_clik_multistart_batched(const sapamd_model *m, int64_t B, int32_t nstart, int32_t link,
                                   const float *target, const float *q0, const int32_t *mask, double eps,
                                   int32_t max_iterations, double dt, double damp, float *q_out, int32_t *success,
                                   float *err, int32_t *iterations, int32_t *solved, void *stream);

/* CLIK problems of several models in one batch (at most 8 models; one
 * persistent launch per phase over all of them): model i solves B[i] problems for its link links[i], with targets[i]
 * (B[i], 7), optional q0[i] (B[i], nv_i) and masks[i] (nv_i), writing
 * q_out[i], success[i], err[i], iterations[i] (may be NULL) like
 * sapamd_clik_batched. Lanes pull problems from one queue, so early-exiting
 * problems of any model free their lane for the next one (no reference
 * counterpart: the reference solves one problem per call,
 * pinocchio_model.cpp:228-300). The queue counters and hand-over records
 * belong to models[0]. */
int sapamd_clik_hetero_batched(const sapamd_model *const *models, int32_t nmodels, const int64_t *B,
                               const int32_t *links, const float *const *targets, const float *const *q0,
                               const int32_t *const *masks, double eps, int32_t max_iterations, double dt,
                               double damp, float *const *q_out, int32_t *const *success, float *const *err,
                               int32_t *const *iterations, void *stream);

/* ---------------------------------------------------------------- adjoints (FP32, batched)
 * No reference counterpart: the reference has no derivatives. Reverse mode of
 * the batched kinematics above, for training losses through the model. */
/* Vector-Jacobian product of link poses and RNEA torques: given upstream
 * gradients gposes (B, L, 7) and/or gtau (B, nv) (either may be NULL; gtau
 * needs qd/qdd), writes dq (B, nv) and, when gtau is given, dqd and dqdd
 * (B, nv). The forward pass is recomputed inside the kernel. */
int sapamd_kin_vjp_batched(const sapamd_model *m, int64_t B, const float *q, const float *qd, const float *qdd,
                           const float *gposes, const float *gtau, float *dq, float *dqd, float *dqdd, void *stream);
/* Fused loss and gradient: loss_b = sum_i (tau_i - tau*_i)^2 + |p_ee - t|^2 with
 * tau = RNEA(q, qd, qdd) and p_ee the position of user link ee_link; tau_star
 * or p_target may be NULL to drop that term. Writes loss (B) and dq, dqd, dqdd
 * (B, nv); tau (B, nv) may be non-NULL to also receive the torques. */
int sapamd_pose_torque_loss_batched(const sapamd_model *m, int64_t B, const float *q, const float *qd,
                                    const float *qdd, const float *tau_star, int32_t ee_link, const float *p_target,
                                    float *loss, float *tau, float *dq, float *dqd, float *dqdd, void *stream);

/* ---------------------------------------------------------------- rollouts (FP32, batched)
 * Semi-implicit Euler with ABA forward dynamics: qdd_t = ABA(q_t, qd_t, tau_t),
 * qd_{t+1} = qd_t + dt qdd_t, q_{t+1} = q_t + dt qd_{t+1}. No reference
 * counterpart (the reference exposes single ABA calls,
 * pinocchio_model.cpp:221-226). Trees of <= 32 joints under the lanes-per-joint
 * kernels run a whole call in one launch (sapamd_set_rollout_launch), other
 * models one launch per step; either way everything is on `stream`, so a
 * caller may capture a whole call in a HIP graph, and nothing is allocated.
 * tau is (B, T, nv). */
/* One step in place on q, qd (B, nv) with torques tau_t rows at tau_stride. */
int sapamd_rollout_step(const sapamd_model *m, int64_t B, float dt, float *q, float *qd, const float *tau_t,
                        int64_t tau_stride, void *stream);
/* T steps in place on q, qd; the pre-step state (q | qd) of steps 0, K, 2K, ...
 * is saved to ckpt (B, ceil(T / K), 2 nv) for the backward pass. */
int sapamd_rollout_forward(const sapamd_model *m, int64_t B, int32_t T, int32_t K, float dt, float *q, float *qd,
                           const float *tau, float *ckpt, void *stream);
/* Backward through T steps from the checkpoints: on entry aq, aqd (B, nv) hold
 * dL/dq_T and dL/dqd_T, on exit dL/dq_0 and dL/dqd_0; gtau (B, T, nv) receives
 * dL/dtau. Segments of K steps are recomputed from their checkpoint into
 * seg (B, K, 2 nv) using q_tmp, qd_tmp (B, nv) as scratch. */
int sapamd_rollout_backward(const sapamd_model *m, int64_t B, int32_t T, int32_t K, float dt, const float *tau,
                            const float *ckpt, float *seg, float *q_tmp, float *qd_tmp, float *aq, float *aqd,
                            float *gtau, void *stream);
/* Rollouts of several models in ONE launch (at most 8 models, each a tree of
 * 1 to 32 joints under the lanes-per-joint kernels, all on one GPU; same T, K
 * and dt): model i steps B[i] environments in place on q[i], qd[i] (B[i], nv_i)
 * under tau[i] (B[i], T, nv_i), checkpoints into ckpt[i] (may be NULL, as may
 * the ckpt array) like sapamd_rollout_forward. Each environment gets the
 * same bits as in its own model's sapamd_rollout_forward (FUSED), in that
 * model's dynamics precision. No reference counterpart (the reference steps
 * one articulation per call, pinocchio_model.cpp:221-226). */
int sapamd_rollout_forward_hetero(const sapamd_model *const *models, int32_t nmodels, const int64_t *B, int32_t T,
                                  int32_t K, float dt, float *const *q, float *const *qd, const float *const *tau,
                                  float *const *ckpt, void *stream);
/* The adjoints of those rollouts in ONE launch, per model as
 * sapamd_rollout_backward (FUSED): aq[i], aqd[i] (B[i], nv_i) hold dL/dq_T,
 * dL/dqd_T on entry and dL/dq_0, dL/dqd_0 on exit; gtau[i] (B[i], T, nv_i)
 * receives dL/dtau; seg[i] (B[i], K, 2 nv_i) is scratch. */
int sapamd_rollout_backward_hetero(const sapamd_model *const *models, int32_t nmodels, const int64_t *B, int32_t T,
                                   int32_t K, float dt, const float *const *tau, const float *const *ckpt,
                                   float *const *seg, float *const *aq, float *const *aqd, float *const *gtau,
                                   void *stream);

/* ---------------------------------------------------------------- forward-dynamics precision
 * Batched forward dynamics (sapamd_aba_batched, rollouts) runs FP32 ABA unless
 * the model's mass matrix is ill-conditioned: at set_joint_order the library
 * estimates cond(M) on the host (FP64 CRBA on 64 seeded configurations within
 * the joint limits) and, above 1e4 (where FP32 ABA can no longer hold 1e-3
 * relative), switches to FP64 CRBA + RNEA bias + LTDL factorization (FP32 in
 * and out). The reference computes ABA in FP64 always
 * (pinocchio_model.cpp:221-226); this keeps its accuracy where FP32 cannot. */
enum { SAPAMD_DYNAMICS_AUTO = 0, SAPAMD_DYNAMICS_FP32 = 1, SAPAMD_DYNAMICS_FP64 = 2 };
int sapamd_set_dynamics_precision(sapamd_model *m, int32_t mode);
/* Effective precision (FP32 or FP64) and the cond(M) estimate. */
int sapamd_get_dynamics_precision(const sapamd_model *m, int32_t *mode, double *cond_estimate);

/* How sapamd_rollout_forward / sapamd_rollout_backward launch for trees of
 * <= 32 joints under the lanes-per-joint kernels: FUSED (default) runs all T
 * steps in one kernel with the state in registers (the adjoint recomputes each
 * K-step segment inside the kernel); PER_STEP issues one launch per step as
 * sapamd_rollout_step does. Same arithmetic either way. No reference
 * counterpart: the reference steps its simulator one call at a time
 * (pinocchio_model.cpp:221-226 for one forward-dynamics evaluation). */
enum { SAPAMD_ROLLOUT_FUSED = 0, SAPAMD_ROLLOUT_PER_STEP = 1 };
int sapamd_set_rollout_launch(sapamd_model *m, int32_t mode);

#ifdef __cplusplus
}
#endif
#endif /* SAPIEN_AMD_H */
