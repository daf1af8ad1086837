"""GPU: parity at the benchmark's workload sizes.

* Gradients (8192 environments of chain7 and of arm9, and one 4096 + 4096
  batch of both): loss, dL/dq, dL/dqd, dL/dqdd of
  loss = |p_ee - FK(q*)|^2 + |RNEA(q, qd, qdd) - tau*|^2 against the FP64
  oracle's exact derivatives (dual-number RNEA derivatives and the position
  Jacobian of the end-effector link) on a 512-environment sample, relative
  error <= 1e-3 per environment; every other environment by batch
  independence (a row's bits do not depend on the batch it is launched in).
* Rollouts (4096 split 4:3:2:1, 128 steps, dt = 1/240, checkpoint every 16):
  a 64-rollout sample of finite final states per tree against the FP64
  oracle; every rollout whose GPU final state is not finite also leaves the
  FP64 oracle's range (non-finite or |x| > 1e6) on the same inputs.
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

from oracle.urdf_model import OracleModel  # noqa: E402
from parity import rel_err  # noqa: E402
from sapien_amd import synthetic as S  # noqa: E402

DEV = "cuda:0"
GRAD_TOL = 1e-3


def _pair(rob):
    import sapien_amd as sa

    m = sa.PinocchioModel(rob.urdf, [0.0, 0.0, -9.81])
    m.set_joint_order(rob.joint_names)
    m.set_link_order(rob.link_names)
    o = OracleModel(rob.urdf)
    o.set_joint_order(rob.joint_names)
    o.set_link_order(rob.link_names)
    return m, o


def _t(x):
    return torch.tensor(np.asarray(x), dtype=torch.float32, device=DEV)


def _loss_and_grads_fp64(o, ee, q, qd, qdd, tau_star, target):
    """Exact FP64 loss and gradients of |p_ee - t|^2 + |RNEA - tau*|^2."""
    r = o.rnea(q, qd, qdd) - tau_star
    P = o.link_poses(q)[:, ee, :3]
    e = P - target
    loss = (r * r).sum(1) + (e * e).sum(1)
    Dq, Dqd, Dqdd = o.rnea_derivs(q, qd, qdd)
    Jp = o.link_pos_jac(q, ee)  # (B, 3, nv) = d p_ee / dq
    dq = 2 * np.einsum("bi,bik->bk", r, Dq) + 2 * np.einsum("bi,bik->bk", e, Jp)
    return loss, dq, 2 * np.einsum("bi,bik->bk", r, Dqd), 2 * np.einsum("bi,bik->bk", r, Dqdd)


def _gradient_batch(m, o, rob, B):
    """Inputs of the benchmark's gradient workload (targets from the FP64 FK)."""
    q, qd, qdd, tau_star, q_star = S.gradient_workload(rob, B)
    target = o.link_poses(q_star.astype(np.float64))[:, rob.ee_link, :3].astype(np.float32)
    return q, qd, qdd, tau_star, target


def _check_sample(o, rob, res, inputs, idx):
    loss, tau, dq, dqd, dqdd = (x.cpu().numpy() for x in res)
    q, qd, qdd, tau_star, target = (x[idx].astype(np.float64) for x in inputs)
    ref = _loss_and_grads_fp64(o, rob.ee_link, q, qd, qdd, tau_star, target)
    assert np.allclose(loss[idx], ref[0], rtol=1e-4, atol=1e-6)
    assert rel_err(tau[idx], o.rnea(q, qd, qdd)) <= 1e-4
    errs = [rel_err(g[idx], r) for g, r in zip((dq, dqd, dqdd), ref[1:])]
    print(f"{rob.name}: grad rel err dq {errs[0]:.2e} dqd {errs[1]:.2e} dqdd {errs[2]:.2e}")
    assert max(errs) <= GRAD_TOL


@pytest.mark.parametrize("tree", ["chain7", "arm9"])
def test_gradients_at_workload_size(tree):
    rob = S.robot(tree)
    m, o = _pair(rob)
    B = 8192
    inputs = _gradient_batch(m, o, rob, B)
    q, qd, qdd, tau_star, target = (_t(x) for x in inputs)
    res = m.pose_torque_loss_batched(q, qd, qdd, tau_star, rob.ee_link, target)
    torch.cuda.synchronize()
    sample = np.random.default_rng(1700).choice(B, 512, replace=False)
    _check_sample(o, rob, res, inputs, sample)
    # batch independence: the sample launched alone gives the same bits
    idx = torch.as_tensor(sample, device=DEV)
    alone = m.pose_torque_loss_batched(q[idx], qd[idx], qdd[idx], tau_star[idx], rob.ee_link, target[idx])
    for a, b in zip(alone, res):
        assert torch.equal(a, b[idx])
    # and the first half as its own launch (the 4096-environment half of the mixed batch)
    half = m.pose_torque_loss_batched(q[:4096], qd[:4096], qdd[:4096], tau_star[:4096], rob.ee_link,
                                      target[:4096])
    for a, b in zip(half, res):
        assert torch.equal(a, b[:4096])


def test_gradients_mixed_4096_4096():
    """One batch of 4096 chain7 + 4096 arm9 environments, as the benchmark
    times it (both trees' loss + gradient launches back to back on one stream)."""
    outs = []
    for tree in ("chain7", "arm9"):
        rob = S.robot(tree)
        m, o = _pair(rob)
        inputs = [x[:4096] for x in _gradient_batch(m, o, rob, 8192)]
        outs.append((rob, m, o, inputs, m.pose_torque_loss_batched(*(_t(x) for x in inputs[:4]), rob.ee_link,
                                                                    _t(inputs[4]))))
    torch.cuda.synchronize()
    for k, (rob, m, o, inputs, res) in enumerate(outs):
        _check_sample(o, rob, res, inputs, np.random.default_rng(1701 + k).choice(4096, 512, replace=False))


# ---------------------------------------------------------------- rollouts
def _blown_up(q, qd):
    x = np.concatenate([q, qd], axis=1)
    return ~np.isfinite(x).all(1) | (np.abs(np.nan_to_num(x, nan=np.inf)) > 1e6).any(1)


# FP32 rollouts over 128 steps: the per-step rounding (2^-24 relative) is
# carried through the trajectory like a perturbation of the state, so the
# bound per rollout is 2e-4 plus 32 times how far the FP64 oracle itself moves
# under an input perturbation of that size (sens, measured per rollout below;
# medians 1e-7 to 9e-7).
ROLL_TOL = 2e-4


@pytest.mark.parametrize("k,tree", list(enumerate(t for t, _ in S.MIX)))
def test_rollouts_at_workload_size(k, tree):
    rob = S.robot(tree)
    m, o = _pair(rob)
    B = S.split_mix(4096)[k]
    q0, qd0, tau = S.rollout_workload(rob, B, 1005 + k)
    qT, qdT, ckpt = m.rollout_batched(_t(q0), _t(qd0), _t(tau), S.ROLLOUT_DT, S.ROLLOUT_CKPT)
    qT, qdT = qT.cpu().numpy(), qdT.cpu().numpy()
    finite = np.isfinite(np.concatenate([qT, qdT], 1)).all(1)
    q64, qd64, tau64 = (x.astype(np.float64) for x in (q0, qd0, tau))
    # every rollout the GPU leaves non-finite also leaves the FP64 oracle's range
    bad = np.flatnonzero(~finite)
    if bad.size:
        rq, rqd = o.rollout(q64[bad], qd64[bad], tau64[bad], S.ROLLOUT_DT)
        print(f"{tree}: {bad.size} of {B} rollouts not finite; oracle blown up for {_blown_up(rq, rqd).sum()}")
        assert _blown_up(rq, rqd).all()
    # a 64-rollout sample of the finite ones against the oracle
    pick = np.random.default_rng(1702 + k).choice(np.flatnonzero(finite), min(64, int(finite.sum())), replace=False)
    rq, rqd = o.rollout(q64[pick], qd64[pick], tau64[pick], S.ROLLOUT_DT)
    assert not _blown_up(rq, rqd).any()
    # the oracle's own sensitivity: the same rollouts from inputs moved by 2^-24 relative
    rng = np.random.default_rng(1703 + k)
    pert = [x * (1 + 2.0 ** -24 * rng.choice([-1.0, 1.0], size=x.shape)) for x in (q64[pick], qd64[pick], tau64[pick])]
    sq, sqd = o.rollout(*pert, S.ROLLOUT_DT)
    ref = np.concatenate([rq, rqd], 1)
    got = np.concatenate([qT[pick], qdT[pick]], 1)
    scale = np.abs(ref).max(1)
    err = np.abs(got - ref).max(1) / scale
    sens = np.abs(np.concatenate([sq, sqd], 1) - ref).max(1) / scale
    print(f"{tree}: {finite.sum()}/{B} finite; sample rel err max {err.max():.2e} median {np.median(err):.2e}; "
          f"oracle sensitivity max {sens.max():.2e} median {np.median(sens):.2e}")
    assert (err <= ROLL_TOL + 32 * sens).all(), (err.max(), sens[err.argmax()])
