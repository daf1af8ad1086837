"""Multi-start batched inverse kinematics.

The motion-planning use of ``compute_inverse_kinematics`` (mplib's
``Planner.IK(target_pose, init_qpos, n_init_qpos=20)``,
docs/source/tutorial/motion_planning/inverse_kinematics.rst:7-19): try the
given start, then up to ``n_init_qpos`` random starts within the joint limits,
and keep the first that converges. Here every (target, start) pair is one CLIK
problem of a single persistent launch; starts are taken in order and a start
is skipped once a lower start of its target has converged, so the answer is
the lowest converged start, as the sequential loop would return.
"""
from __future__ import annotations

import torch


def inverse_kinematics_multistart(model, link_index, poses, initial_qpos=None, n_init_qpos=20, active_qmask=None,
                                  eps=1e-4, max_iterations=1000, dt=0.1, damp=1e-6, seed=0, stream=None):
    """poses (B, 7) p | q wxyz -> (qpos (B, nv), success (B,), error (B, 6), start (B,)).

    start is the index of the start that produced the answer (0 = initial_qpos,
    which defaults to the neutral configuration); for targets that no start
    reached, the start with the smallest final error (its success is False).
    """
    dev = torch.device("cuda", torch.cuda.current_device())
    poses = torch.as_tensor(poses, dtype=torch.float32, device=dev)
    if poses.dim() == 1:
        poses = poses[None]
    B, nv = poses.shape[0], model.nv
    n_start = 1 + int(n_init_qpos)
    lo, hi = (torch.as_tensor(x, dtype=torch.float32, device=dev) for x in model.joint_limits())
    gen = torch.Generator(device=dev)
    gen.manual_seed(int(seed))
    q0 = torch.empty(n_start, B, nv, device=dev)
    q0[0] = 0.0 if initial_qpos is None else torch.as_tensor(initial_qpos, dtype=torch.float32, device=dev)
    if n_start > 1:
        q0[1:] = lo + (hi - lo) * torch.rand(n_start - 1, B, nv, generator=gen, device=dev)
    Q, S, E, I, solved = model._clik_multistart_raw(link_index, poses, q0.reshape(-1, nv), n_start, active_qmask,
                                                    eps, max_iterations, dt, damp, stream)
    Q, S, E, I = Q.view(n_start, B, nv), S.view(n_start, B), E.view(n_start, B, 6), I.view(n_start, B)
    fallback = E.norm(dim=2).argmin(dim=0)  # skipped starts carry error 1e30
    start = torch.where(solved < n_start, solved.long(), fallback)
    t = torch.arange(B, device=dev)
    return Q[start, t], S[start, t].bool(), E[start, t], start
