"""torch.autograd.Function wrappers over the hand-written adjoint kernels.

* ``link_poses(model, q)``            -> (B, L, 7) poses; backward = FK pose VJP kernel
* ``inverse_dynamics(model, q, qd, qdd)`` -> (B, nv) torques; backward = RNEA VJP kernel
* ``pose_torque_loss(model, q, qd, qdd, tau_target, ee_link, p_target)``
      -> (B,) loss_b = |tau - tau*|^2 + |p_ee - t|^2; the forward kernel returns
      the loss and its gradients in one pass, backward only scales them.
* ``rollout(model, q0, qd0, tau, dt, checkpoint_every=16)`` -> final (q_T, qd_T)
      of semi-implicit Euler with ABA over tau (B, T, nv); the forward keeps the
      state every K steps, backward recomputes each K-step segment from its
      checkpoint and runs the step adjoint kernel backwards through it.

Backward passes recompute the forward inside the adjoint kernel from the saved
inputs (no intermediates are stored between forward and backward). Kernels run
in FP32; gradients are returned in the dtype of the inputs. The Jacobian output
of the batched API has no adjoint (its derivative is second order).
"""
from __future__ import annotations

import torch


def _like(x, ref):
    return x.to(dtype=ref.dtype) if x is not None and ref is not None and x.dtype != ref.dtype else x


class _LinkPoses(torch.autograd.Function):
    @staticmethod
    def forward(ctx, model, q):
        ctx.model = model
        ctx.save_for_backward(q)
        return model.compute_forward_kinematics_batched(q.detach()).to(q.dtype)

    @staticmethod
    def backward(ctx, gposes):
        (q,) = ctx.saved_tensors
        dq, _, _ = ctx.model.kinematics_vjp_batched(q.detach(), grad_poses=gposes.detach().contiguous())
        return None, _like(dq, q)


class _InverseDynamics(torch.autograd.Function):
    @staticmethod
    def forward(ctx, model, q, qd, qdd):
        ctx.model = model
        ctx.save_for_backward(q, qd, qdd)
        return model.compute_inverse_dynamics_batched(q.detach(), qd.detach(), qdd.detach()).to(q.dtype)

    @staticmethod
    def backward(ctx, gtau):
        q, qd, qdd = ctx.saved_tensors
        dq, dqd, dqdd = ctx.model.kinematics_vjp_batched(q.detach(), qd.detach(), qdd.detach(),
                                                         grad_tau=gtau.detach().contiguous())
        return None, _like(dq, q), _like(dqd, qd), _like(dqdd, qdd)


class _PoseTorqueLoss(torch.autograd.Function):
    @staticmethod
    def forward(ctx, model, q, qd, qdd, tau_target, ee_link, p_target):
        loss, tau, dq, dqd, dqdd = model.pose_torque_loss_batched(
            q.detach(), None if qd is None else qd.detach(), None if qdd is None else qdd.detach(),
            None if tau_target is None else tau_target.detach(), ee_link,
            None if p_target is None else p_target.detach())
        ctx.save_for_backward(dq, dqd, dqdd)
        ctx.dtypes = (q.dtype, None if qd is None else qd.dtype, None if qdd is None else qdd.dtype)
        return loss.to(q.dtype)

    @staticmethod
    def backward(ctx, gloss):
        dq, dqd, dqdd = ctx.saved_tensors
        s = gloss.detach().to(torch.float32)[:, None]
        out = [None, (dq * s).to(ctx.dtypes[0])]
        out.append(None if ctx.dtypes[1] is None else (dqd * s).to(ctx.dtypes[1]))
        out.append(None if ctx.dtypes[2] is None else (dqdd * s).to(ctx.dtypes[2]))
        return tuple(out) + (None, None, None)


class _Rollout(torch.autograd.Function):
    @staticmethod
    def forward(ctx, model, q0, qd0, tau, dt, K):
        qT, qdT, ckpt = model.rollout_batched(q0.detach(), qd0.detach(), tau.detach(), dt, K)
        ctx.model, ctx.dt, ctx.K = model, dt, K
        ctx.dtypes = (q0.dtype, qd0.dtype, tau.dtype)
        ctx.save_for_backward(tau.detach().to(torch.float32).contiguous(), ckpt)
        return qT.to(q0.dtype), qdT.to(qd0.dtype)

    @staticmethod
    def backward(ctx, gq, gqd):
        tau, ckpt = ctx.saved_tensors
        gq = torch.zeros_like(tau[:, 0]) if gq is None else gq.detach().to(torch.float32).contiguous()
        gqd = torch.zeros_like(tau[:, 0]) if gqd is None else gqd.detach().to(torch.float32).contiguous()
        dq0, dqd0, dtau = ctx.model.rollout_backward_batched(tau, ckpt, ctx.dt, gq, gqd, ctx.K)
        return None, dq0.to(ctx.dtypes[0]), dqd0.to(ctx.dtypes[1]), dtau.to(ctx.dtypes[2]), None, None


def link_poses(model, q):
    return _LinkPoses.apply(model, q)


def inverse_dynamics(model, q, qd, qdd):
    return _InverseDynamics.apply(model, q, qd, qdd)


def pose_torque_loss(model, q, qd=None, qdd=None, tau_target=None, ee_link=0, p_target=None):
    return _PoseTorqueLoss.apply(model, q, qd, qdd, tau_target, ee_link, p_target)


def rollout(model, q0, qd0, tau, dt, checkpoint_every=16):
    return _Rollout.apply(model, q0, qd0, tau, dt, checkpoint_every)
