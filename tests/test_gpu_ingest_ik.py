"""GPU: what the reference's factory feeds the path, and its IK fixture.

* URDF text exactly as SAPIEN's exporter writes it (dummy links, x-axis joint
  frames, FP32 values printed by str, continuous joints with +-1e5 limits, a
  fixed world root; tests/exporter_style.py, after
  python/py_package/wrapper/urdf_exporter.py:21-152) through the native parser
  and the HIP kernels: the drop-in FP64 path and the batched FP32 path against
  the FP64 oracle reading the same text; the oracle on the exported text
  against the oracle on the source URDF within the FP32 quantization.
* The reference's own IK fixture: panda link 9 to
  Pose([-0.01, -0.39, 0.96], [0.21, 0.73, 0.41, 0.5]) (manualtest/urdf.py:32-33;
  the quaternion is not unit length and is used as given, as Eigen does),
  through the drop-in and the batched path against oracle.clik.
* Continuous joints in IK output: atan2-wrapped on and off the chain
  (posP2S over the whole configuration, pinocchio_model.cpp:87,299).
"""
import os
import xml.etree.ElementTree as ET

import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

from exporter_style import export  # noqa: E402
from oracle.urdf_model import OracleModel  # noqa: E402
from parity import POSE_TOL_M, POSE_TOL_RAD, REL_TOL, pose_errors, rel_err  # noqa: E402
from sapien_amd import synthetic as S  # noqa: E402

DEV = "cuda:0"
GOLD = os.path.join(os.path.dirname(__file__), "golden")


def _t(x):
    return torch.tensor(np.asarray(x), dtype=torch.float32, device=DEV)


def _models(text, jn, ln):
    import sapien_amd as sa

    m = sa.PinocchioModel(text, [0.0, 0.0, -9.81])
    m.set_joint_order(jn)
    m.set_link_order(ln)
    o = OracleModel(text)
    o.set_joint_order(jn)
    o.set_link_order(ln)
    return m, o


def _source(name):
    """(urdf, movable joint names, link names, q sampler) of a synthetic tree or a golden robot."""
    if name in S.ROBOTS:
        rob = S.robot(name)
        return rob.urdf, rob.joint_names, rob.link_names, lambda n, seed: S.samples(rob, n, seed)
    text = open(os.path.join(GOLD, name)).read()
    root = ET.fromstring(text)
    jn = [j.get("name") for j in root.findall("joint") if j.get("type") != "fixed"]
    ln = [l.get("name") for l in root.findall("link")]

    def sampler(n, seed):
        rng = np.random.default_rng(seed)
        return [rng.uniform(-1, 1, (n, len(jn))), rng.normal(size=(n, len(jn))), rng.normal(0, 2, (n, len(jn))),
                rng.normal(0, 3, (n, len(jn)))]
    return text, jn, ln, sampler


def _exported(name, seed=11):
    text, jn, ln, sampler = _source(name)
    etext, jmap, lmap = export(text, seed)
    return text, jn, ln, etext, [jmap[j] for j in jn], [lmap[l] for l in ln], sampler


@pytest.mark.parametrize("name", ["chain7", "torso30", "panda.urdf", "robots/xarm7.urdf"])
def test_exporter_urdf_through_the_native_path(name):
    text, jn, ln, etext, ejn, eln, sampler = _exported(name)
    assert etext.startswith("<?xml") and "link_dummy_" in etext
    m, o = _models(etext, ejn, eln)
    q, qd, qdd, tau = sampler(256, 2100)
    # drop-in FP64 per call against the oracle on the same text
    for b in range(3):
        m.compute_forward_kinematics(q[b])
        Pr = o.link_poses(q[b])[0]
        for l in range(len(eln)):
            pose = m.get_link_pose(l)
            assert np.allclose(pose.p, Pr[l, :3], atol=2e-7)
            assert np.allclose(pose.q, Pr[l, 3:], atol=2e-7) or np.allclose(pose.q, -Pr[l, 3:], atol=2e-7)
        m.compute_full_jacobian(q[b])
        for l in range(len(eln)):
            assert np.allclose(m.get_link_jacobian(l, False), o.link_jacobian(q[b], l, False)[0], atol=1e-11)
        assert np.allclose(m.compute_generalized_mass_matrix(q[b]), o.crba(q[b])[0], atol=1e-11)
        assert np.allclose(m.compute_inverse_dynamics(q[b], qd[b], qdd[b]), o.rnea(q[b], qd[b], qdd[b])[0],
                           atol=1e-9)
    # batched FP32 against the FP64 oracle
    P, J, T = (x.cpu().numpy() for x in m.compute_fk_jacobian_rnea_batched(_t(q), _t(qd), _t(qdd)))
    pe, ae = pose_errors(P, o.link_poses(q))
    assert pe <= POSE_TOL_M and ae <= POSE_TOL_RAD
    assert rel_err(J, o.jacobian_world(q)) <= REL_TOL
    assert rel_err(T, o.rnea(q, qd, qdd)) <= REL_TOL
    M = m.compute_generalized_mass_matrix_batched(_t(q)).cpu().numpy()
    assert rel_err(M, o.crba(q)) <= REL_TOL
    # the exported text is the source mechanism to FP32 quantization
    os_ = OracleModel(text)
    os_.set_joint_order(jn)
    os_.set_link_order(ln)
    assert np.abs(o.link_poses(q)[..., :3] - os_.link_poses(q)[..., :3]).max() < 1e-6
    assert rel_err(o.jacobian_world(q), os_.jacobian_world(q)) < 1e-5
    if "<inertial" in text:  # (panda.urdf has none: the export then holds placeholder inertias)
        assert rel_err(o.rnea(q, qd, qdd), os_.rnea(q, qd, qdd)) < 1e-5


PANDA_TARGET = np.array([-0.01, -0.39, 0.96, 0.21, 0.73, 0.41, 0.5])  # manualtest/urdf.py:32


def test_panda_reference_ik_fixture():
    """compute_inverse_kinematics(9, Pose([-0.01, -0.39, 0.96], [0.21, 0.73,
    0.41, 0.5])) on the panda loaded as create_pinocchio_model loads it (the
    exporter's URDF, links in SAPIEN's index order: link 9 = panda_hand), from
    the default neutral start. The first iterates equal the oracle's; the
    start is near-singular (J J^T + 1e-6 I), which amplifies rounding
    differences from step to step, so the full runs are compared by their
    outcome: both converge, and the answer reaches the target in the oracle's
    FK (log6 error below eps), for the drop-in and the batched path."""
    import sapien_amd as sa

    text, jn, ln, etext, ejn, eln, _ = _exported("panda.urdf")
    assert ln[9] == "panda_hand"
    m, o = _models(etext, ejn, eln)
    t32 = PANDA_TARGET.astype(np.float32)
    t64 = t32.astype(np.float64)
    pose = sa.Pose(t32[:3], t32[3:])
    for k in (1, 2, 3):
        qo, ok, err = m.compute_inverse_kinematics(9, pose, max_iterations=k)
        qr, okr, errr, _ = o.clik(9, t64, max_iterations=k)
        dq, de = np.abs(qo - qr).max(), np.abs(err - errr).max()
        print(f"panda iterate {k}: |dq| {dq:.2e} |derr| {de:.2e}")
        assert dq < 1e-7 and de < 1e-7, k
    qr, okr, errr, itr = o.clik(9, t64)
    qo, ok, err = m.compute_inverse_kinematics(9, pose)
    Q, okb, errb, itb = (x.cpu().numpy() for x in m.compute_inverse_kinematics_batched(9, _t(t32[None])))
    print(f"panda: oracle success {okr} |err| {np.linalg.norm(errr):.3e} after {itr}; drop-in success {ok} "
          f"|err| {np.linalg.norm(err):.3e}; batched success {okb[0]} |err| {np.linalg.norm(errb[0]):.3e} "
          f"after {itb[0]}")
    assert okr and ok and okb[0]
    # the answers' error in the oracle's own FP64 evaluation (zero iterations from them)
    for qq, e, tol in ((qo, err, 1e-4), (Q[0].astype(np.float64), errb[0], 1.5e-4)):
        _, ok0, e0, it0 = o.clik(9, t64, q0=qq, max_iterations=0)
        assert it0 == 0 and np.linalg.norm(e0) < tol and np.linalg.norm(e) < 1e-4


def _with_continuous(rob, joints):
    """The robot with the given joints made continuous (no limits)."""
    root = ET.fromstring(rob.urdf)
    for j in root.findall("joint"):
        if j.get("name") in joints:
            j.set("type", "continuous")
            lim = j.find("limit")
            if lim is not None:
                j.remove(lim)
    return ET.tostring(root, encoding="unicode")


def test_ik_wraps_continuous_joints_on_and_off_the_chain():
    """chain7 with joint_2 continuous (on the chain of link 7), and torso30's
    wheels (continuous, off the chain of its arm link 16): every continuous dof
    of the returned configuration is atan2(sin q, cos q), on the drop-in and
    both batched kernel families, as the oracle returns it. A wheel started at
    q0 = 4.0 comes back as 4.0 - 2 pi."""
    import sapien_amd as sa

    # on the chain
    rob = S.chain7()
    text = _with_continuous(rob, {"joint_2"})
    m, o = _models(text, rob.joint_names, rob.link_names)
    assert m.nq == 8
    q = S.samples(rob, 8, 2101)[0]
    q[:, 1] += 2 * np.pi * np.array([1, -1, 2, -2, 1, -1, 3, 0])  # the same pose, the angle out of (-pi, pi]
    tgt = o.link_poses(q)[:, 7]
    q0 = q + 0.05
    for b in range(4):
        t32 = tgt[b].astype(np.float32)
        qo, ok, err = m.compute_inverse_kinematics(7, sa.Pose(t32[:3], t32[3:]), initial_qpos=q0[b])
        qr, okr, errr, _ = o.clik(7, t32.astype(np.float64), q0=q0[b])
        assert ok and okr and -np.pi < qo[1] <= np.pi
        assert np.allclose(qo, qr, atol=1e-6)
    # batched: the same local solution as the oracle from the same start (the
    # chain is redundant, so not necessarily the sampled q), angle wrapped
    ref = np.stack([o.clik(7, tgt[b].astype(np.float32).astype(np.float64),
                           q0=q0[b].astype(np.float32).astype(np.float64))[0] for b in range(len(q))])
    assert (ref[:, 1] > -np.pi).all() and (ref[:, 1] <= np.pi).all()
    for variant in ("auto", "thread_per_env"):
        sa.set_kernel_variant(m, variant)
        Q, ok, err, it = (x.cpu().numpy() for x in m.compute_inverse_kinematics_batched(7, _t(tgt), _t(q0)))
        assert ok.all(), variant
        assert (Q[:, 1] > -np.pi - 1e-6).all() and (Q[:, 1] <= np.pi + 1e-6).all(), variant
        d = (Q - ref + np.pi) % (2 * np.pi) - np.pi
        print(f"continuous on chain, {variant}: |Q - oracle| {np.abs(d).max():.2e}")
        assert np.abs(d).max() < 2e-3, variant
    # off the chain: torso30's wheels, with q0 = 4.0 on every wheel
    rob = S.torso30()
    m, o = _models(rob.urdf, rob.joint_names, rob.link_names)
    wheels = np.flatnonzero(rob.continuous)
    assert len(wheels) == 6
    q = S.samples(rob, 8, 2102)[0]
    tgt = o.link_poses(q)[:, rob.ee_link]
    q0 = q + 0.05
    q0[:, wheels] = 4.0
    t32 = tgt[0].astype(np.float32)
    qo, ok, err = m.compute_inverse_kinematics(rob.ee_link, sa.Pose(t32[:3], t32[3:]), initial_qpos=q0[0])
    qr, okr, errr, _ = o.clik(rob.ee_link, t32.astype(np.float64), q0=q0[0])
    assert np.allclose(qo[wheels], 4.0 - 2 * np.pi, atol=1e-12) and np.allclose(qr[wheels], 4.0 - 2 * np.pi)
    assert ok == okr and np.allclose(qo, qr, atol=1e-6)
    for variant in ("auto", "thread_per_env"):
        sa.set_kernel_variant(m, variant)
        Q = m.compute_inverse_kinematics_batched(rob.ee_link, _t(tgt), _t(q0))[0].cpu().numpy()
        assert np.allclose(Q[:, wheels], 4.0 - 2 * np.pi, atol=1e-6), variant
