"""GPU: hand-written adjoint kernels against the FP64 oracle.

Tolerance: per environment max|d - d_ref| / max|d_ref| <= 1e-3 (FP32 kernels
against exact FP64 dual-number derivatives and FP64 central differences).
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


def _fk_vjp_fd(o, q, gP, h=1e-6):
    out = np.zeros_like(q)
    for b in range(q.shape[0]):
        P0 = o.link_poses(q[b])[0]
        for k in range(q.shape[1]):
            e = np.zeros(q.shape[1])
            e[k] = h
            Pp, Pm = o.link_poses(q[b] + e)[0], o.link_poses(q[b] - e)[0]
            for X in (Pp, Pm):
                flip = (X[:, 3:] * P0[:, 3:]).sum(1) < 0
                X[flip, 3:] *= -1
            out[b, k] = ((Pp - Pm) * gP[b]).sum() / (2 * h)
    return out


@pytest.mark.parametrize("tree", ["chain7", "arm9", "quad18", "torso30"])
def test_rnea_vjp_matches_exact_derivatives(tree):
    import sapien_amd as sa

    rob = S.robot(tree)
    m, o = _pair(rob)
    q, qd, qdd, _ = S.samples(rob, 64, 1400)
    g = np.random.default_rng(1).normal(size=q.shape)
    Dq, Dqd, Dqdd = o.rnea_derivs(q, qd, qdd)
    variants = ["auto", "thread_per_env"]
    for v in variants:
        sa.set_kernel_variant(m, v)
        dq, dqd, dqdd = m.kinematics_vjp_batched(_t(q), _t(qd), _t(qdd), grad_tau=_t(g))
        assert rel_err(dq.cpu().numpy(), np.einsum("bi,bik->bk", g, Dq)) <= GRAD_TOL, v
        assert rel_err(dqd.cpu().numpy(), np.einsum("bi,bik->bk", g, Dqd)) <= GRAD_TOL, v
        assert rel_err(dqdd.cpu().numpy(), np.einsum("bi,bik->bk", g, Dqdd)) <= GRAD_TOL, v


@pytest.mark.parametrize("tree", ["chain7", "arm9", "torso30"])
def test_fk_pose_vjp_matches_central_differences(tree):
    import sapien_amd as sa

    rob = S.robot(tree)
    m, o = _pair(rob)
    q = S.samples(rob, 8, 1401)[0]
    gP = np.random.default_rng(2).normal(size=(8, rob.nlinks, 7))
    ref = _fk_vjp_fd(o, q, gP)
    for v in ("auto", "thread_per_env"):
        sa.set_kernel_variant(m, v)
        dq = m.kinematics_vjp_batched(_t(q), grad_poses=_t(gP))[0].cpu().numpy()
        assert rel_err(dq, ref) <= GRAD_TOL, v


@pytest.mark.parametrize("tree", ["chain7", "arm9"])
def test_pose_torque_loss_gradients(tree):
    """Config: loss = |p_ee - FK(q*)|^2 + |tau_RNEA - tau*|^2, gradients against FP64 central differences."""
    import sapien_amd as sa

    rob = S.robot(tree)
    m, o = _pair(rob)
    q, qd, qdd, tau_star = S.samples(rob, 16, 1402)
    qs = S.samples(rob, 16, 1403)[0]
    ee = rob.ee_link
    t = o.link_poses(qs)[:, ee, :3]

    def loss64(q_, qd_, qdd_):
        return ((o.rnea(q_, qd_, qdd_) - tau_star) ** 2).sum(1) + ((o.link_poses(q_)[:, ee, :3] - t) ** 2).sum(1)

    ref_loss = loss64(q, qd, qdd)
    h = 1e-6
    fd = []
    for x in range(3):
        d = np.zeros_like(q)
        for k in range(rob.nv):
            e = np.zeros_like(q)
            e[:, k] = h
            args_p = [q + e * (x == 0), qd + e * (x == 1), qdd + e * (x == 2)]
            args_m = [q - e * (x == 0), qd - e * (x == 1), qdd - e * (x == 2)]
            d[:, k] = (loss64(*args_p) - loss64(*args_m)) / (2 * h)
        fd.append(d)
    for v in ("auto", "thread_per_env"):
        sa.set_kernel_variant(m, v)
        loss, tau, dq, dqd, dqdd = m.pose_torque_loss_batched(_t(q), _t(qd), _t(qdd), _t(tau_star), ee, _t(t))
        assert np.allclose(loss.cpu().numpy(), ref_loss, rtol=1e-4), v
        assert rel_err(tau.cpu().numpy(), o.rnea(q, qd, qdd)) <= 1e-4
        for got, ref in zip((dq, dqd, dqdd), fd):
            assert rel_err(got.cpu().numpy(), ref) <= GRAD_TOL, v


def test_autograd_functions():
    from sapien_amd import autograd as A

    rob = S.chain7()
    m, o = _pair(rob)
    q, qd, qdd, tau_star = S.samples(rob, 32, 1404)
    qt = _t(q).requires_grad_()
    qdt = _t(qd).requires_grad_()
    qddt = _t(qdd).requires_grad_()
    tau = A.inverse_dynamics(m, qt, qdt, qddt)
    P = A.link_poses(m, qt)
    (tau.square().sum() + P[:, 7, :3].sum()).backward()
    g = 2 * o.rnea(q, qd, qdd)
    Dq, Dqd, Dqdd = o.rnea_derivs(q, qd, qdd)
    assert rel_err(qdt.grad.cpu().numpy(), np.einsum("bi,bik->bk", g, Dqd)) <= GRAD_TOL
    assert rel_err(qddt.grad.cpu().numpy(), np.einsum("bi,bik->bk", g, Dqdd)) <= GRAD_TOL
    pos_jac = o.link_pos_jac(q, 7)  # (B, 3, nv)
    ref_q = np.einsum("bi,bik->bk", g, Dq) + pos_jac.sum(axis=1)
    assert rel_err(qt.grad.cpu().numpy(), ref_q) <= GRAD_TOL
    # fused loss op: backward scales the in-kernel gradients by the upstream gradient
    qt2 = _t(q).requires_grad_()
    t = _t(o.link_poses(S.samples(rob, 32, 1405)[0])[:, 7, :3])
    loss = A.pose_torque_loss(m, qt2, _t(qd), _t(qdd), _t(tau_star), 7, t)
    (3.0 * loss.sum()).backward()
    _, _, dq, _, _ = m.pose_torque_loss_batched(_t(q), _t(qd), _t(qdd), _t(tau_star), 7, t)
    assert torch.allclose(qt2.grad, 3.0 * dq)


def test_chain_vjp_every_shape_with_permuted_orders():
    """Every lanes-per-joint shape of the chain adjoint kernel (including the
    interleaved 8 x 1 one) with a permuted user joint and link order and a
    ragged batch: RNEA VJP against the exact FP64 derivatives, pose VJP and the
    fused loss against the thread-per-environment kernel."""
    import sapien_amd as sa
    from sapien_amd import tune

    rob = S.chain7()
    rng = np.random.default_rng(1406)
    jn = [rob.joint_names[i] for i in rng.permutation(len(rob.joint_names))]
    lperm = list(rng.permutation(len(rob.link_names)))
    ln = [rob.link_names[i] for i in lperm]
    ee = lperm.index(rob.ee_link)
    m = sa.PinocchioModel(rob.urdf, [0.0, 0.0, -9.81])
    m.set_joint_order(jn)
    m.set_link_order(ln)
    o = OracleModel(rob.urdf)
    o.set_joint_order(jn)
    o.set_link_order(ln)
    B = 300
    q, qd, qdd, tau_star = S.samples(rob, B, 1407)
    g = rng.normal(size=q.shape)
    gP = rng.normal(size=(B, rob.nlinks, 7))
    t = _t(o.link_poses(S.samples(rob, B, 1408)[0])[:, ee, :3])
    Dq, Dqd, Dqdd = o.rnea_derivs(q, qd, qdd)
    refs = [np.einsum("bi,bik->bk", g, D) for D in (Dq, Dqd, Dqdd)]
    sa.set_kernel_variant(m, "thread_per_env")
    ref_pose = m.kinematics_vjp_batched(_t(q), grad_poses=_t(gP))[0].cpu().numpy()
    ref_loss = [x.cpu().numpy() for x in m.pose_torque_loss_batched(_t(q), _t(qd), _t(qdd), _t(tau_star), ee, t)]
    sa.set_kernel_variant(m, "lanes_per_joint")
    try:
        for G, C in tune.shapes():
            if G * C < rob.nv:
                continue
            tune.set_shape(m, 1 << 40, G, C)
            got = m.kinematics_vjp_batched(_t(q), _t(qd), _t(qdd), grad_tau=_t(g))
            for x, ref in zip(got, refs):
                assert rel_err(x.cpu().numpy(), ref) <= GRAD_TOL, (G, C)
            dq = m.kinematics_vjp_batched(_t(q), grad_poses=_t(gP))[0].cpu().numpy()
            assert rel_err(dq, ref_pose) <= GRAD_TOL, (G, C)
            res = m.pose_torque_loss_batched(_t(q), _t(qd), _t(qdd), _t(tau_star), ee, t)
            assert np.allclose(res[0].cpu().numpy(), ref_loss[0], rtol=1e-4), (G, C)
            for x, ref in zip(res[1:], ref_loss[1:]):
                assert rel_err(x.cpu().numpy(), ref) <= GRAD_TOL, (G, C)
    finally:
        tune.set_shape(m, 0, 0, 0)
