"""GPU parity: HIP kernels (through the C ABI) against the FP64 oracle.

* drop-in per-call methods (FP64 on the GPU, reference semantics) match the
  oracle to ~1e-9 for every tree;
* batched FP32 methods at the workload sizes (7-dof chain at B = 16384, the
  mixed 8192 batch split 3277/2458/1638/819) meet the FP32 tolerances written
  in tests/parity.py;
* edge cases: B = 0, B = 1, B not a multiple of the wave size, CPU inputs,
  wrong shapes.
"""
import math
import os

import numpy as np
import pytest

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu

from oracle.urdf_model import OracleModel  # noqa: E402
from parity import (POSE_TOL_M, POSE_TOL_RAD, REL_TOL, link_state_reference, pose_errors, quat_to_R,  # noqa: E402
                    rel_err)
from sapien_amd import synthetic as S  # noqa: E402

TREES = list(S.ROBOTS)
MIX = {"chain7": 3277, "arm9": 2458, "quad18": 1638, "torso30": 819}
DEV = "cuda:0"


def _pair(rob, urdf=None):
    import sapien_amd as sa

    m = sa.PinocchioModel(urdf or rob.urdf, [0.0, 0.0, -9.81])
    m.set_joint_order(rob.joint_names)
    m.set_link_order(rob.link_names)
    o = OracleModel(urdf or rob.urdf)
    o.set_joint_order(rob.joint_names)
    o.set_link_order(rob.link_names)
    return m, o


def _t(x):
    return torch.tensor(np.asarray(x), dtype=torch.float32, device=DEV)


# ----------------------------------------------------------------- drop-in (FP64 per call)
@pytest.mark.parametrize("tree", TREES)
def test_dropin_fp64_matches_oracle(tree):
    rob = S.robot(tree)
    m, o = _pair(rob)
    q, qd, qdd, tau = S.samples(rob, 3, 2001)
    for b in range(3):
        m.compute_forward_kinematics(q[b])
        Pr = o.link_poses(q[b])[0]
        for l in range(rob.nlinks):
            pose = m.get_link_pose(l)
            assert np.allclose(pose.p, Pr[l, :3], atol=2e-7)
            assert np.allclose(pose.q, Pr[l, 3:], atol=2e-7) or np.allclose(pose.q, -Pr[l, 3:], atol=2e-7)
        m.compute_full_jacobian(q[b])
        for l in range(rob.nlinks):
            for local in (False, True):
                Jr = o.link_jacobian(q[b], l, local)[0]
                assert np.allclose(m.get_link_jacobian(l, local), Jr, atol=1e-11), (l, local)
            assert np.allclose(m.compute_single_link_local_jacobian(q[b], l), o.link_jacobian(q[b], l, True)[0],
                               atol=1e-11)
        assert np.allclose(m.compute_generalized_mass_matrix(q[b]), o.crba(q[b])[0], atol=1e-11)
        assert np.allclose(m.compute_inverse_dynamics(q[b], qd[b], qdd[b]), o.rnea(q[b], qd[b], qdd[b])[0],
                           atol=1e-9)
        assert np.allclose(m.compute_forward_dynamics(q[b], qd[b], tau[b]), o.aba(q[b], qd[b], tau[b])[0],
                           rtol=1e-8, atol=1e-8)
        assert np.allclose(m.compute_coriolis_matrix(q[b], qd[b]), o.coriolis(q[b], qd[b])[0], atol=1e-9)


@pytest.mark.parametrize("n", [40, 60])
def test_dropin_fp64_long_chains_match_oracle(n):
    """Per-call methods on chains past 32 joints (the 48- and 64-joint
    one-environment kernels: per-joint work and output entries spread over
    the block's threads, per-joint arrays in LDS) against the FP64 oracle."""
    rob = S.chain(n, seed=60 + n)
    m, o = _pair(rob)
    q, qd, qdd, tau = S.samples(rob, 2, 2101 + n)
    for b in range(2):
        m.compute_forward_kinematics(q[b])
        Pr = o.link_poses(q[b])[0]
        for l in range(rob.nlinks):
            pose = m.get_link_pose(l)
            assert np.allclose(pose.p, Pr[l, :3], atol=2e-6)
            assert np.allclose(pose.q, Pr[l, 3:], atol=2e-7) or np.allclose(pose.q, -Pr[l, 3:], atol=2e-7)
        m.compute_full_jacobian(q[b])
        for l in (0, rob.nlinks // 2, rob.nlinks - 1):
            for local in (False, True):
                assert rel_err(m.get_link_jacobian(l, local), o.link_jacobian(q[b], l, local)[0]) <= 1e-12
        assert rel_err(m.compute_generalized_mass_matrix(q[b]), o.crba(q[b])[0]) <= 1e-12
        assert rel_err(m.compute_inverse_dynamics(q[b], qd[b], qdd[b]), o.rnea(q[b], qd[b], qdd[b])[0]) <= 1e-12
        assert rel_err(m.compute_coriolis_matrix(q[b], qd[b]), o.coriolis(q[b], qd[b])[0]) <= 1e-12
        assert rel_err(m.compute_forward_dynamics(q[b], qd[b], tau[b]), o.aba(q[b], qd[b], tau[b])[0]) <= 1e-8
    # IK from 0.05 rad of a solution: both converge, to the same configuration
    import sapien_amd as sa

    ee = rob.nlinks - 1
    qs = S.samples(rob, 1, 2201 + n)[0][0]
    q0 = qs + np.random.default_rng(2202 + n).normal(scale=0.05, size=qs.shape)
    t7 = o.link_poses(qs)[0, ee].astype(np.float32).astype(np.float64)
    qo, ok, err = m.compute_inverse_kinematics(ee, sa.Pose(t7[:3], t7[3:]), initial_qpos=q0)
    qr, okr, errr, _ = o.clik(ee, t7, q0=q0)
    assert ok and okr
    assert np.allclose(qo, qr, atol=1e-6)


def test_dropin_models_interleaved():
    """Per-call methods of two models called alternately: each model reads and
    writes its own pinned per-call buffer, so results and caches stay apart."""
    ra, rb = S.robot("chain7"), S.robot("torso30")
    (ma, oa), (mb, ob) = _pair(ra), _pair(rb)
    qa, qda, qdda, _ = S.samples(ra, 4, 2301)
    qb, qdb, qddb, _ = S.samples(rb, 4, 2302)
    for k in range(4):
        ma.compute_forward_kinematics(qa[k])
        tb = mb.compute_inverse_dynamics(qb[k], qdb[k], qddb[k])
        mb.compute_forward_kinematics(qb[k])
        ta = ma.compute_inverse_dynamics(qa[k], qda[k], qdda[k])
        Pa, Pb = oa.link_poses(qa[k])[0], ob.link_poses(qb[k])[0]
        for m, P, n in ((ma, Pa, ra.nlinks), (mb, Pb, rb.nlinks)):
            for l in range(n):
                pose = m.get_link_pose(l)
                assert np.allclose(pose.p, P[l, :3], atol=2e-7)
        assert np.allclose(ta, oa.rnea(qa[k], qda[k], qdda[k])[0], atol=1e-9)
        assert np.allclose(tb, ob.rnea(qb[k], qdb[k], qddb[k])[0], atol=1e-9)


def test_dropin_movo_fixture_reference_semantics(golden_dir):
    """unittest/test_physx/test_articulation.py:75-97 semantics on the movo fixture."""
    import xml.etree.ElementTree as ET

    import sapien_amd as sa
    from bruteforce import BruteRobot

    text = open(os.path.join(golden_dir, "movo_simple.urdf")).read()
    root = ET.fromstring(text)
    jn = [j.get("name") for j in root.findall("joint") if j.get("type") != "fixed"]
    ln = [l.get("name") for l in root.findall("link")]
    m = sa.PinocchioModel(text, [0, 0, -9.81])
    m.set_joint_order(jn)
    m.set_link_order(ln)
    o = OracleModel(text)
    o.set_joint_order(jn)
    o.set_link_order(ln)
    q = np.array([0.1, 0.1, 0.1, 0.2, 0.2, 0.2, 0.2, 0.2, 0.2, 0.2, 0.3, 0.4, 0.5])
    m.compute_forward_kinematics(q)
    Pr = o.link_poses(q)[0]
    for l in range(len(ln)):
        pose = m.get_link_pose(l)
        assert np.allclose(pose.p, Pr[l, :3], rtol=1e-5, atol=1e-5)
        assert np.allclose(pose.q, Pr[l, 3:], rtol=1e-5, atol=1e-5) or np.allclose(pose.q, -Pr[l, 3:], rtol=1e-5,
                                                                                   atol=1e-5)
    g = m.compute_inverse_dynamics(q, np.zeros_like(q), np.zeros_like(q))
    assert np.allclose(g, BruteRobot(text, jn).gravity_torque(q), rtol=1e-5, atol=1e-5)


def test_dropin_ik_matches_oracle():
    rob = S.chain7()
    m, o = _pair(rob)
    q = S.samples(rob, 4, 2002)[0]
    P = o.link_poses(q)
    import sapien_amd as sa

    for b in range(4):
        tgt = sa.Pose(P[b, 7, :3].astype(np.float32), P[b, 7, 3:].astype(np.float32))
        qo, ok, err = m.compute_inverse_kinematics(7, tgt, initial_qpos=q[b] + 0.3)
        t7 = np.r_[tgt.p, tgt.q].astype(np.float64)
        qr, okr, errr, it = o.clik(7, t7, q0=q[b] + 0.3)
        assert ok == okr
        assert np.allclose(qo, qr, atol=1e-6)
        assert np.allclose(err, errr, atol=1e-8)
        # default initial_qpos = neutral; the mask freezes joint 2 at its neutral 0,
        # so the target comes from a configuration with q[2] = 0 (reachable)
        qm = q[b].copy()
        qm[2] = 0.0
        Pm = o.link_poses(qm)[0, 7]
        tgt = sa.Pose(Pm[:3].astype(np.float32), Pm[3:].astype(np.float32))
        t7 = np.r_[tgt.p, tgt.q].astype(np.float64)
        mask = [1, 1, 0, 1, 1, 1, 1]
        # the iteration itself: the first step from neutral gives the same iterate
        # (later steps from this far start go through near-singular J J^T + 1e-6 I,
        # which amplifies rounding differences by ~1e6 per step)
        for k in (1,):
            qo, ok, err = m.compute_inverse_kinematics(7, tgt, active_qmask=mask, max_iterations=k)
            qr, okr, errr, it = o.clik(7, t7, mask=mask, max_iterations=k)
            assert qo[2] == 0.0 and np.allclose(qo, qr, atol=1e-9) and np.allclose(err, errr, atol=1e-9)
        # full runs from a far start: the early steps branch chaotically, so the
        # two runs may converge to different (equally valid) solutions or end on
        # either side of eps; a converged result must reach the target
        qo, ok, err = m.compute_inverse_kinematics(7, tgt, active_qmask=mask)
        assert qo[2] == 0.0
        if ok:
            assert np.linalg.norm(err) < 1e-4
            reached = o.link_poses(qo)[0, 7]
            assert np.abs(reached[:3] - t7[:3]).max() < 1e-4


def test_dropin_errors_on_gpu():
    rob = S.chain7()
    m, _ = _pair(rob)
    with pytest.raises(RuntimeError, match="posS2P failed"):
        m.compute_inverse_dynamics(np.zeros(6), np.zeros(6), np.zeros(6))
    with pytest.raises(RuntimeError, match="link index out of bound"):
        m.compute_single_link_local_jacobian(np.zeros(7), 8)


# ----------------------------------------------------------------- batched FP32
def _check_kin(m, o, q, qd, qdd):
    P, J, T = m.compute_fk_jacobian_rnea_batched(_t(q), _t(qd), _t(qdd))
    torch.cuda.synchronize()
    pos, ang = pose_errors(P.cpu().numpy(), o.link_poses(q))
    assert pos <= POSE_TOL_M and ang <= POSE_TOL_RAD, (pos, ang)
    assert rel_err(J.cpu().numpy(), o.jacobian_world(q)) <= REL_TOL
    assert rel_err(T.cpu().numpy(), o.rnea(q, qd, qdd)) <= REL_TOL


def test_batched_chain7_metric_size():
    rob = S.chain7()
    m, o = _pair(rob)
    q, qd, qdd, tau = S.samples(rob, 16384, 1001)
    _check_kin(m, o, q, qd, qdd)
    # the separate entry points agree with the fused one bit for bit
    P, J, T = m.compute_fk_jacobian_rnea_batched(_t(q), _t(qd), _t(qdd))
    assert torch.equal(m.compute_forward_kinematics_batched(_t(q)), P)
    assert torch.equal(m.compute_full_jacobian_batched(_t(q)), J)
    assert torch.equal(m.compute_inverse_dynamics_batched(_t(q), _t(qd), _t(qdd)), T)
    M = m.compute_generalized_mass_matrix_batched(_t(q)).cpu().numpy()
    assert rel_err(M, o.crba(q)) <= REL_TOL
    A = m.compute_forward_dynamics_batched(_t(q), _t(qd), _t(tau)).cpu().numpy()
    Ar = o.aba(q, qd, tau)
    # ABA in FP32 carries cond(M) * eps: bounded against the FP64 result by 1e-3 relative
    assert rel_err(A, Ar) <= 1e-3


def test_batched_chain7_large_batches_are_tiled_copies():
    """Batch sizes of the throughput sweep (262,144 and 1,048,576 environments)
    cannot be compared against the FP64 oracle in seconds; an environment's
    result does not depend on its position in the batch, so the 16384-sample
    batch tiled 16x and 64x must give the 16384 results tiled, bit for bit, in
    the interleaved 8 x 1 and the 4 x 2 shapes."""
    from sapien_amd import tune

    rob = S.chain7()
    m, o = _pair(rob)
    q, qd, qdd, _ = (_t(x) for x in S.samples(rob, 16384, 1001))
    try:
        for G, C in ((8, 1), (4, 2)):
            tune.set_shape(m, 1 << 40, G, C)
            ref = m.compute_fk_jacobian_rnea_batched(q, qd, qdd)
            for reps in (16, 64):
                out = m.compute_fk_jacobian_rnea_batched(q.repeat(reps, 1), qd.repeat(reps, 1), qdd.repeat(reps, 1))
                for a, r in zip(out, ref):
                    assert torch.equal(a, r.repeat(reps, *([1] * (r.dim() - 1)))), (G, C, reps)
                del out
    finally:
        tune.set_shape(m, 0, 0, 0)


@pytest.mark.parametrize("tree", TREES)
def test_batched_mixed_trees(tree):
    rob = S.robot(tree)
    m, o = _pair(rob)
    B = MIX[tree]
    q, qd, qdd, tau = S.samples(rob, B, 1002)
    _check_kin(m, o, q, qd, qdd)
    M = m.compute_generalized_mass_matrix_batched(_t(q)).cpu().numpy()
    assert rel_err(M, o.crba(q)) <= REL_TOL
    Jw, Jl = m.compute_link_jacobians_batched(_t(q[:64]))
    for l in range(rob.nlinks):
        assert rel_err(Jw[:, l].cpu().numpy(), o.link_jacobian(q[:64], l, False) + 0) <= REL_TOL or \
            np.abs(o.link_jacobian(q[:64], l, False)).max() == 0
        assert np.abs(Jl[:, l].cpu().numpy() - o.link_jacobian(q[:64], l, True)).max() <= 1e-4 * max(
            1.0, np.abs(o.link_jacobian(q[:64], l, True)).max())
    C = m.compute_coriolis_matrix_batched(_t(q[:256]), _t(qd[:256])).cpu().numpy()
    assert rel_err(C, o.coriolis(q[:256], qd[:256])) <= 1e-3


def test_batched_edge_cases():
    rob = S.arm9()
    m, o = _pair(rob)
    for B in (1, 65, 127):
        q, qd, qdd, _ = S.samples(rob, B, 1003 + B)
        _check_kin(m, o, q, qd, qdd)
    P, J, T = m.compute_fk_jacobian_rnea_batched(_t(np.zeros((0, 9))), _t(np.zeros((0, 9))), _t(np.zeros((0, 9))))
    assert P.shape == (0, 10, 7) and J.shape == (0, 6, 9) and T.shape == (0, 9)
    # CPU numpy inputs are moved to the model's GPU
    q = S.samples(rob, 5, 1010)[0]
    assert m.compute_forward_kinematics_batched(q).device.type == "cuda"
    with pytest.raises(RuntimeError):
        m.compute_forward_kinematics_batched(_t(np.zeros((4, 8))))


@pytest.mark.parametrize("variant", ["auto", "thread_per_env"])
def test_batched_clik(variant):
    """Batched CLIK on chain7, lanes-per-joint (auto) and thread-per-problem."""
    import sapien_amd as sa

    rob = S.chain7()
    m, o = _pair(rob)
    sa.set_kernel_variant(m, variant)
    B = 4096
    qs = S.samples(rob, B, 1004)[0]
    tgt = o.link_poses(qs)[:, 7]
    Q, ok, err, it = m.compute_inverse_kinematics_batched(7, _t(tgt))
    Q, ok, err, it = Q.cpu().numpy(), ok.cpu().numpy(), err.cpu().numpy(), it.cpu().numpy()
    assert ok.mean() > 0.5
    assert (np.linalg.norm(err[ok], axis=1) < 1e-4).all()
    reached = o.link_poses(Q[ok])[:, 7]
    assert np.abs(reached[:, :3] - tgt[ok, :3]).max() < 2e-4
    # same algorithm as the FP64 per-problem oracle: same outcome on a sample
    agree = 0
    for b in range(32):
        qr, okr, errr, itr = o.clik(7, tgt[b])
        agree += okr == ok[b]
    assert agree >= 28


def test_batched_clik_results_do_not_depend_on_the_batch():
    """The lanes-per-joint CLIK runs every problem for up to 128 iterations in
    one launch and hands the ones still running to a second launch through a
    record buffer (stuck-looking ones first). A problem's result must not
    depend on what else is in the batch or where its record lands: the same
    512 reachable problems solved alone and interleaved with 512 unreachable
    ones (targets moved 5 m away, which run to the iteration limit) give the
    same bits; the unreachable ones report failure after max_iterations."""
    rob = S.chain7()
    m, o = _pair(rob)
    n = 512
    tgt = o.link_poses(S.samples(rob, n, 1009)[0])[:, 7]
    far = o.link_poses(S.samples(rob, n, 1010)[0])[:, 7].copy()
    far[:, :3] += 5.0
    mixed = np.empty((2 * n, 7))
    mixed[0::2], mixed[1::2] = tgt, far
    alone = [x.cpu().numpy() for x in m.compute_inverse_kinematics_batched(7, _t(tgt))]
    both = [x.cpu().numpy() for x in m.compute_inverse_kinematics_batched(7, _t(mixed))]
    assert (alone[3] > 128).any()  # some reachable problems went through the hand-over
    for a, b in zip(alone, both):
        assert np.array_equal(a, b[0::2])
    Q, ok, err, it = (x[1::2] for x in both)
    assert not ok.any() and (it == 1000).all()
    assert (np.linalg.norm(err, axis=1) > 1.0).all()


# ----------------------------------------------------------------- kernel families agree
@pytest.mark.parametrize("n", [3, 7, 12, 16])
def test_lanes_per_joint_matches_oracle_and_generic(n):
    import sapien_amd as sa

    rob = S.chain(n, seed=40 + n)
    m, o = _pair(rob)
    assert sa.kernel_variant(m) == "lanes_per_joint"
    q, qd, qdd, _ = S.samples(rob, 1000, 1100 + n)
    _check_kin(m, o, q, qd, qdd)
    fast = m.compute_fk_jacobian_rnea_batched(_t(q), _t(qd), _t(qdd))
    sa.set_kernel_variant(m, "thread_per_env")
    assert sa.kernel_variant(m) == "thread_per_env"
    gen = m.compute_fk_jacobian_rnea_batched(_t(q), _t(qd), _t(qdd))
    for a, b in zip(fast, gen):
        assert rel_err(a.cpu().numpy(), b.cpu().numpy().astype(np.float64)) <= REL_TOL
    sa.set_kernel_variant(m, "auto")
    # trees run the tree lanes-per-joint kernel; beyond 32 joints only thread-per-env exists
    m9, _ = _pair(S.arm9())
    assert sa.kernel_variant(m9) == "lanes_per_joint" and not m9.is_chain
    big = S.chain(40, seed=5)
    mb, _ = _pair(big)
    assert sa.kernel_variant(mb) == "thread_per_env"
    with pytest.raises(RuntimeError):
        sa.set_kernel_variant(mb, "lanes_per_joint")


def test_fast_kernel_partial_outputs_and_ragged_batches():
    rob = S.chain7()
    m, o = _pair(rob)
    for B in (1, 5, 31, 33, 1000):
        q, qd, qdd, _ = S.samples(rob, B, 1200 + B)
        _check_kin(m, o, q, qd, qdd)
        P = m.compute_forward_kinematics_batched(_t(q))
        J = m.compute_full_jacobian_batched(_t(q))
        T = m.compute_inverse_dynamics_batched(_t(q), _t(qd), _t(qdd))
        P2, J2, T2 = m.compute_fk_jacobian_rnea_batched(_t(q), _t(qd), _t(qdd))
        assert torch.equal(P, P2) and torch.equal(J, J2) and torch.equal(T, T2)


@pytest.mark.parametrize("n", [7, 12])
def test_every_chain_shape_matches_oracle(n):
    from sapien_amd import tune

    rob = S.chain7() if n == 7 else S.chain(n, seed=3)
    m, o = _pair(rob)
    q, qd, qdd, _ = S.samples(rob, 300, 1300 + n)
    for G, C in tune.shapes():
        if G * C < n:
            continue
        tune.set_shape(m, 1 << 40, G, C)
        assert tune.get_shape(m, 300) == (G, C)
        _check_kin(m, o, q, qd, qdd)
    tune.set_shape(m, 0, 0, 0)  # clear tuned entries


# ----------------------------------------------------------------- trees and heterogeneous launches
@pytest.mark.parametrize("tree", ["arm9", "quad18", "torso30"])
def test_tree_kernel_matches_oracle_and_generic(tree):
    import sapien_amd as sa

    rob = S.robot(tree)
    m, o = _pair(rob)
    assert sa.kernel_variant(m) == "lanes_per_joint" and not m.is_chain
    q, qd, qdd, _ = S.samples(rob, 777, 1500)
    _check_kin(m, o, q, qd, qdd)
    M = m.compute_generalized_mass_matrix_batched(_t(q)).cpu().numpy()
    assert rel_err(M, o.crba(q)) <= REL_TOL
    fast = m.compute_fk_jacobian_rnea_batched(_t(q), _t(qd), _t(qdd)) + (m.compute_generalized_mass_matrix_batched(_t(q)),)
    sa.set_kernel_variant(m, "thread_per_env")
    gen = m.compute_fk_jacobian_rnea_batched(_t(q), _t(qd), _t(qdd)) + (m.compute_generalized_mass_matrix_batched(_t(q)),)
    for a, b in zip(fast, gen):
        assert rel_err(a.cpu().numpy(), b.cpu().numpy().astype(np.float64)) <= REL_TOL


def test_tree_kernel_general_link_lists(golden_dir):
    """movo_simple: several links per joint (fixed frames) takes the general link path."""
    import xml.etree.ElementTree as ET

    import sapien_amd as sa

    text = open(os.path.join(golden_dir, "movo_simple.urdf")).read()
    root = ET.fromstring(text)
    jn = [j.get("name") for j in root.findall("joint") if j.get("type") != "fixed"]
    ln = [l.get("name") for l in root.findall("link")]
    m = sa.PinocchioModel(text, [0, 0, -9.81])
    m.set_joint_order(jn)
    m.set_link_order(ln)
    o = OracleModel(text)
    o.set_joint_order(jn)
    o.set_link_order(ln)
    rng = np.random.default_rng(8)
    q = rng.uniform(-1, 1, size=(200, 13))
    q[:, 0] = rng.uniform(0, 0.4, size=200)
    qd, qdd = rng.normal(size=(200, 13)), rng.normal(size=(200, 13))
    assert sa.kernel_variant(m) == "lanes_per_joint"
    _check_kin(m, o, q, qd, qdd)
    assert rel_err(m.compute_generalized_mass_matrix_batched(_t(q)).cpu().numpy(), o.crba(q)) <= REL_TOL


def test_heterogeneous_batch_one_launch():
    """The mixed 8192 batch (4:3:2:1) of the four trees in ONE kernel launch."""
    import sapien_amd as sa

    pairs, qs, qds, qdds, refs = [], [], [], [], []
    for i, tree in enumerate(["chain7", "arm9", "quad18", "torso30"]):
        rob = S.robot(tree)
        m, o = _pair(rob)
        q, qd, qdd, _ = S.samples(rob, MIX[tree], 1600 + i)
        pairs.append((m, o))
        qs.append(_t(q))
        qds.append(_t(qd))
        qdds.append(_t(qdd))
        refs.append((o.link_poses(q), o.jacobian_world(q), o.rnea(q, qd, qdd), o.crba(q)))
    outs = sa._pinocchio.hetero_batched([p[0] for p in pairs], qs, qds, qdds)
    torch.cuda.synchronize()
    for out, (P, J, T, M) in zip(outs, refs):
        pos, ang = pose_errors(out["poses"].cpu().numpy(), P)
        assert pos <= POSE_TOL_M and ang <= POSE_TOL_RAD
        assert rel_err(out["J"].cpu().numpy(), J) <= REL_TOL
        assert rel_err(out["tau"].cpu().numpy(), T) <= REL_TOL
        assert rel_err(out["M"].cpu().numpy(), M) <= REL_TOL


def test_heterogeneous_batch_does_not_depend_on_model_order():
    """The launch dispatches the costliest trees' blocks first whatever order
    the caller lists the models in: each model's outputs are the same bits in
    either order, and the same as a launch of that model alone."""
    import sapien_amd as sa

    models, ins = [], []
    for i, tree in enumerate(["chain7", "arm9", "quad18", "torso30"]):
        rob = S.robot(tree)
        m, _ = _pair(rob)
        q, qd, qdd, _ = S.samples(rob, 300 + 37 * i, 1650 + i)
        models.append(m)
        ins.append((_t(q), _t(qd), _t(qdd)))

    def run(order):
        outs = sa._pinocchio.hetero_batched([models[k] for k in order], [ins[k][0] for k in order],
                                            [ins[k][1] for k in order], [ins[k][2] for k in order])
        return {k: o for k, o in zip(order, outs)}

    fwd, rev = run([0, 1, 2, 3]), run([3, 2, 1, 0])
    torch.cuda.synchronize()
    for k in range(4):
        alone = run([k])[k]
        for key in ("poses", "J", "tau", "M"):
            assert torch.equal(fwd[k][key], rev[k][key]), (k, key)
            assert torch.equal(fwd[k][key], alone[key]), (k, key)


@pytest.mark.parametrize("tree", ["chain7", "arm9", "quad18", "torso30"])
def test_clik_kernel_families_agree(tree):
    """Lanes-per-joint CLIK (one problem per 8 or 16 lanes) against the
    thread-per-problem kernel and the FP64 oracle. The algorithm is the same up
    to the association order of the FK products and J J^T sums.
    Well-conditioned problems (q0 within 0.05 rad of a solution): four
    iterations agree with each other and with the oracle to rounding.
    Random starts with joint 0 masked: damped least squares (damp 1e-6) near
    singular poses amplifies the FP32 differences of the first error (1e-5,
    largest where the rotation error is near pi) by the conditioning of J J^T,
    so individual trajectories part; the success rates match, and every
    converged problem of either kernel reaches its target with the masked
    joint left at q0."""
    import sapien_amd as sa

    rob = S.robot(tree)
    m, o = _pair(rob)
    link = rob.ee_link
    # well conditioned: four iterations against each other and the oracle
    B = 256
    qs = S.samples(rob, B, 1722)[0]
    q0n = (qs + np.random.default_rng(1723).normal(scale=0.05, size=qs.shape)).astype(np.float32).astype(np.float64)
    t = o.link_poses(qs)[:, link].astype(np.float32).astype(np.float64)
    short = {}
    for variant in ("auto", "thread_per_env"):
        sa.set_kernel_variant(m, variant)
        short[variant] = [x.cpu().numpy() for x in
                          m.compute_inverse_kinematics_batched(link, _t(t), _t(q0n), max_iterations=4)]
    (Ql, okl, el, itl), (Qt, okt, et, itt) = short["auto"], short["thread_per_env"]
    d = np.abs(Ql.astype(np.float64) - Qt).max()
    dor = max(np.abs(Ql[b] - o.clik(link, t[b], q0n[b], max_iterations=4)[0]).max() for b in range(16))
    print(f"{tree}: 4 iterations, lanes vs thread {d:.1e}, lanes vs oracle {dor:.1e}")
    assert d <= 1e-4 and dor <= 1e-4
    assert np.array_equal(okl, okt) and np.array_equal(itl, itt)

    # random starts, joint 0 masked: statistics and correctness
    B = 512
    qs = S.samples(rob, B, 1720)[0]
    q0n = S.samples(rob, B, 1721)[0]
    qs[:, 0] = q0n[:, 0]  # the masked joint already at its target value
    t = _t(o.link_poses(qs)[:, link])
    q0 = _t(q0n)
    mask = torch.ones(rob.nv, dtype=torch.int32, device=DEV)
    mask[0] = 0
    full = {}
    for variant in ("auto", "thread_per_env"):
        sa.set_kernel_variant(m, variant)
        full[variant] = [x.cpu().numpy() for x in m.compute_inverse_kinematics_batched(link, t, q0, mask)]
    fl, ft = full["auto"], full["thread_per_env"]
    print(f"{tree}: success lanes {fl[1].mean():.3f} thread {ft[1].mean():.3f}; "
          f"mean iterations {fl[3].mean():.1f} / {ft[3].mean():.1f}")
    assert abs(fl[1].mean() - ft[1].mean()) <= 0.05
    for Q, ok, err, _ in (fl, ft):
        assert (np.linalg.norm(err[ok], axis=1) < 1e-4).all()
        reached = o.link_poses(Q[ok].astype(np.float64))[:, link, :3]
        assert np.abs(reached - t.cpu().numpy()[ok, :3]).max() < 2e-4
        # the masked joint keeps q0 (wrapped by atan2 when continuous, as posP2S does)
        q00 = q0n[:, 0].astype(np.float32)
        if rob.continuous[0]:
            assert np.allclose(Q[:, 0], np.arctan2(np.sin(q00), np.cos(q00)), atol=1e-6)
        else:
            assert np.array_equal(Q[:, 0], q00)


def test_heterogeneous_clik_one_launch():
    """CLIK problems of the four trees (4:3:2:1) in ONE persistent launch:
    bit-identical to each model's own batched CLIK (same per-problem code),
    and every converged problem reaches its target."""
    import sapien_amd as sa

    models, links, tgts = [], [], []
    for i, (tree, nprob) in enumerate(zip(["chain7", "arm9", "quad18", "torso30"], [800, 600, 400, 200])):
        rob = S.robot(tree)
        m, o = _pair(rob)
        qs = S.samples(rob, nprob, 1700 + i)[0]
        models.append((m, o, rob))
        links.append(rob.ee_link)
        tgts.append(_t(o.link_poses(qs)[:, rob.ee_link]))
    outs = sa.clik_hetero_batched([x[0] for x in models], links, tgts)
    for (m, o, rob), link, t, (Q, ok, err, it) in zip(models, links, tgts, outs):
        Qr, okr, errr, itr = m.compute_inverse_kinematics_batched(link, t)
        assert torch.equal(Q, Qr) and torch.equal(ok, okr) and torch.equal(err, errr) and torch.equal(it, itr), rob.name
        okn = ok.cpu().numpy()
        assert okn.mean() > 0.5, (rob.name, okn.mean())
        assert (np.linalg.norm(err.cpu().numpy()[okn], axis=1) < 1e-4).all()
        reached = o.link_poses(Q.cpu().numpy()[okn])[:, link, :3]
        assert np.abs(reached - t.cpu().numpy()[okn, :3]).max() < 2e-4


def test_fused_launch_graph_replay_matches_eager():
    """The C-ABI fused launch captured in a HIP graph (two buffer sets) and
    replayed three times, with eager work between replays: every replay equals
    the eager launch bit for bit."""
    import sapien_amd as sa

    rob = S.chain7()
    m, _ = _pair(rob)
    launch = sa.native_fk_jac_rnea(m)
    B, L, n = 3000, rob.nlinks, rob.nv
    sets = []
    for k in range(2):
        q, qd, qdd, _ = S.samples(rob, B, 1800 + k)
        sets.append([_t(q), _t(qd), _t(qdd), torch.empty(B, L, 7, device=DEV), torch.empty(B, 6, n, device=DEV),
                     torch.empty(B, n, device=DEV)])
    ref = [m.compute_fk_jacobian_rnea_batched(s[0], s[1], s[2]) for s in sets]
    g = torch.cuda.CUDAGraph()
    cap = torch.cuda.Stream()
    cap.wait_stream(torch.cuda.current_stream())
    with torch.cuda.graph(g, stream=cap):
        cs = torch.cuda.current_stream().cuda_stream
        for s in sets:
            launch(B, *s, cs)
    torch.cuda.current_stream().wait_stream(cap)
    for rep in range(3):
        for s in sets:
            for t in s[3:]:
                t.fill_(float("nan"))
        g.replay()
        torch.cuda.synchronize()
        for s, r in zip(sets, ref):
            assert all(torch.equal(a, b) for a, b in zip(s[3:], r)), rep
        m.compute_fk_jacobian_rnea_batched(sets[0][0], sets[0][1], sets[0][2])  # eager work between replays
        torch.cuda.synchronize()


@pytest.mark.parametrize("tree", ["chain7", "arm9"])
def test_timed_launch_records_the_kernel(tree):
    """sapamd_fk_jac_rnea_batched_timed: the same outputs as the plain launch,
    and the two events bracket the kernel (positive, shorter than the host
    time around launch + synchronize); the thread-per-env variant refuses."""
    import time

    import sapien_amd as sa

    rob = S.robot(tree)
    m, _ = _pair(rob)
    B, L, n = 4096, rob.nlinks, rob.nv
    q, qd, qdd, _ = (_t(x) for x in S.samples(rob, B, 1810))
    out = [torch.full((B, L, 7), float("nan"), device=DEV), torch.full((B, 6, n), float("nan"), device=DEV),
           torch.full((B, n), float("nan"), device=DEV)]
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    e1.record()
    torch.cuda.synchronize()
    timed = sa.native_fk_jac_rnea_timed(m)
    t0 = time.perf_counter()
    timed(B, q, qd, qdd, *out, e0.cuda_event, e1.cuda_event, torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    host_ms = (time.perf_counter() - t0) * 1e3
    ref = m.compute_fk_jacobian_rnea_batched(q, qd, qdd)
    assert all(torch.equal(a, b) for a, b in zip(out, ref))
    ms = e0.elapsed_time(e1)
    print(f"{tree}: kernel {ms * 1e3:.1f} us, host {host_ms * 1e3:.1f} us")
    assert 0 < ms < host_ms
    sa.set_kernel_variant(m, "thread_per_env")
    with pytest.raises(Exception, match="lanes-per-joint"):
        timed(B, q, qd, qdd, *out, e0.cuda_event, e1.cuda_event, torch.cuda.current_stream().cuda_stream)


@pytest.mark.parametrize("scale", [16.0, 200.0])
def test_large_joint_angles(scale):
    """Joint angles far outside the sampled range (both sides of the 16 rad
    switch between the hardware sine and the full-range routine) keep the pose,
    Jacobian and torque tolerances, in the chain and tree kernels."""
    for tree in ("chain7", "arm9"):
        rob = S.robot(tree)
        m, o = _pair(rob)
        q, qd, qdd, _ = S.samples(rob, 2048, 1900)
        # revolute joints only (prismatic fingers keep their range), at
        # FP32-representable angles so that the oracle sees the kernel's inputs
        rev = np.asarray(rob.upper) > 1.0
        big = np.random.default_rng(5).uniform(-scale, scale, size=q.shape).astype(np.float32).astype(np.float64)
        q = np.where(rev[None, :], big, q)
        _check_kin(m, o, q, qd, qdd)


@pytest.mark.parametrize("tree", TREES)
def test_passive_force_batched(tree):
    """compute_passive_force_batched = PhysxArticulation.compute_passive_force
    semantics (src/physx/articulation.cpp:400-418; the reference's test equates
    it with RNEA(q, 0, 0) at rest, unittest/test_physx/test_articulation.py:90-97):
    g(q) + C(q, qd) qd, each term optional; inputs may be padded rows."""
    import sapien_amd as sa

    rob = S.robot(tree)
    m, o = _pair(rob)
    B, n = 1000, rob.nv
    q, qd, _, _ = S.samples(rob, B, 1950)
    z = np.zeros_like(q)
    g_ref, full_ref = o.rnea(q, z, z), o.rnea(q, qd, z)
    # padded [B, max_dof] state buffers, the model reads the first nv columns
    qbuf = torch.full((B, n + 5), 7.0, device=DEV)
    qdbuf = torch.full((B, n + 3), 7.0, device=DEV)
    qbuf[:, :n] = _t(q)
    qdbuf[:, :n] = _t(qd)
    for v in ("auto", "thread_per_env"):
        sa.set_kernel_variant(m, v)
        both = m.compute_passive_force_batched(qbuf[:, :n], qdbuf[:, :n]).cpu().numpy()
        grav = m.compute_passive_force_batched(_t(q), coriolis_and_centrifugal=False).cpu().numpy()
        cori = m.compute_passive_force_batched(_t(q), _t(qd), gravity=False).cpu().numpy()
        none = m.compute_passive_force_batched(_t(q), gravity=False, coriolis_and_centrifugal=False).cpu().numpy()
        assert rel_err(both, full_ref) <= REL_TOL, v
        assert rel_err(grav, g_ref) <= REL_TOL, v
        assert rel_err(cori, full_ref - g_ref) <= REL_TOL, v
        assert np.array_equal(none, np.zeros_like(none))
        # at rest the passive force is the inverse dynamics of zero motion
        rest = m.compute_passive_force_batched(_t(q), torch.zeros(B, n, device=DEV))
        assert torch.equal(rest, m.compute_inverse_dynamics_batched(_t(q), torch.zeros(B, n, device=DEV),
                                                                    torch.zeros(B, n, device=DEV)))


@pytest.mark.parametrize("tree", TREES)
def test_padded_simulator_buffers(tree):
    """The fused kernel reads a padded [n_art, max_dof] joint state in place and
    writes p | q wxyz into the first 7 floats of each link of a
    [n_art, max_links, 13] body-state buffer (the layout of
    src/physx/physx_system.cpp:1048-1101), leaving every other float alone."""
    import sapien_amd as sa

    rob = S.robot(tree)
    m, o = _pair(rob)
    B, n, L = 700, rob.nv, rob.nlinks
    q, qd, qdd, _ = S.samples(rob, B, 1951)
    max_dof, max_links = n + 4, L + 3
    state = torch.full((4, B, max_dof), -3.0, device=DEV)  # qpos, qvel, qacc, qf blocks
    state[0, :, :n], state[1, :, :n], state[2, :, :n] = _t(q), _t(qd), _t(qdd)
    body = torch.full((B, max_links, 13), -5.0, device=DEV)
    J = torch.empty(B, 6, n, device=DEV)
    for v in ("auto", "thread_per_env"):
        sa.set_kernel_variant(m, v)
        body.fill_(-5.0)
        state[3].fill_(-3.0)
        m.compute_fk_jacobian_rnea_into(state[0, :, :n], state[1, :, :n], state[2, :, :n], poses=body, jacobian=J,
                                        qf=state[3, :, :n])
        P, Jr, T = m.compute_fk_jacobian_rnea_batched(_t(q), _t(qd), _t(qdd))
        assert torch.equal(body[:, :L, :7], P) and torch.equal(J, Jr) and torch.equal(state[3, :, :n], T), v
        assert bool((body[:, :, 7:] == -5.0).all()) and bool((body[:, L:, :] == -5.0).all())
        assert bool((state[3, :, n:] == -3.0).all())
        pos, ang = pose_errors(body[:, :L, :7].cpu().numpy(), o.link_poses(q))
        assert pos <= POSE_TOL_M and ang <= POSE_TOL_RAD


@pytest.mark.parametrize("B", [16389, 1, 9])
def test_chain_outputs_staged_through_lds(B):
    """The streaming chain shapes write each wavefront's outputs as contiguous
    runs through LDS: a partial last wavefront (B not a multiple of 8 or 4),
    and output arrays 16-byte aligned (float4 runs) or 4 bytes off (float
    runs), give the same bits as the direct row stores of the strided entry
    point, and nothing outside the outputs is written."""
    import sapien_amd as sa
    from sapien_amd import tune

    rob = S.chain7()
    m, o = _pair(rob)
    n, L = rob.nv, rob.nlinks
    q, qd, qdd, _ = S.samples(rob, B, 1960)
    qt, qdt, qddt = _t(q), _t(qd), _t(qdd)
    launch = sa.native_fk_jac_rnea(m)
    cs = torch.cuda.current_stream().cuda_stream
    for G, C in ((8, 1), (8, 2), (16, 1)):
        tune.set_shape(m, 1 << 40, G, C)
        # reference: the strided entry point with padded rows takes the direct stores
        state = torch.zeros(B, n + 1, device=DEV)
        body = torch.full((B, L + 1, 13), -5.0, device=DEV)
        Jr = torch.empty(B, 6, n, device=DEV)
        state[:, :n] = qt
        m.compute_fk_jacobian_rnea_into(state[:, :n], qdt, qddt, poses=body, jacobian=Jr, qf=None)
        for off in (0, 1):  # element offset of every output array
            Pbuf = torch.full((B * L * 7 + 8,), -7.0, device=DEV)
            Jbuf = torch.full((B * 6 * n + 8,), -7.0, device=DEV)
            Tbuf = torch.full((B * n + 8,), -7.0, device=DEV)
            P = Pbuf[off:off + B * L * 7].view(B, L, 7)
            J = Jbuf[off:off + B * 6 * n].view(B, 6, n)
            T = Tbuf[off:off + B * n].view(B, n)
            launch(B, qt, qdt, qddt, P, J, T, cs)
            torch.cuda.synchronize()
            assert torch.equal(P, body[:, :L, :7]) and torch.equal(J, Jr), (G, C, off)
            for buf, used in ((Pbuf, B * L * 7), (Jbuf, B * 6 * n), (Tbuf, B * n)):
                assert bool((buf[:off] == -7.0).all()) and bool((buf[off + used:] == -7.0).all()), (G, C, off)
            _, _, Tp = m.compute_fk_jacobian_rnea_batched(qt, qdt, qddt)
            assert torch.equal(T, Tp), (G, C, off)
            assert rel_err(T.cpu().numpy(), o.rnea(q, qd, qdd)) <= REL_TOL
    tune.set_shape(m, 0, 0, 0)


@pytest.mark.parametrize("tree", TREES)
def test_simulator_link_state(tree):
    """sapamd_link_state_strided on a padded [n_art, max_links, 13] body-state
    buffer (src/physx/physx_system.cpp:1048-1051) whose row 0 holds a moving,
    non-identity root pose and velocity: every link's row receives the world
    pose root * (model pose) and the centre-of-mass and angular velocities,
    against the oracle's root-frame link poses and WORLD link Jacobians J q_dot
    composed with the root motion; padding rows and a padded joint state are
    left alone; without qvel only the poses change."""
    rob = S.robot(tree)
    m, o = _pair(rob)
    B, n, L = 600, rob.nv, rob.nlinks
    q, qd, _, _ = S.samples(rob, B, 1952)
    rng = np.random.default_rng(1953)
    max_dof, max_links = n + 3, L + 2
    state = torch.full((2, B, max_dof), -3.0, device=DEV)
    state[0, :, :n], state[1, :, :n] = _t(q), _t(qd)
    root_q = rng.normal(size=(B, 4))
    root_q /= np.linalg.norm(root_q, axis=1, keepdims=True)
    root = np.concatenate([rng.uniform(-2, 2, (B, 3)), root_q, rng.normal(size=(B, 3)), rng.normal(size=(B, 3))], 1)
    body = torch.full((B, max_links, 13), -5.0, device=DEV)
    body[:, 0] = _t(root)
    m.compute_link_state_into(state[0, :, :n], state[1, :, :n], body)
    got = body.cpu().numpy().astype(np.float64)
    assert bool((body[:, L:] == -5.0).all()) and bool((state[:, :, n:] == -3.0).all())
    pw, Rw, vc, ww = link_state_reference(o, rob.urdf, rob.link_names, q, qd, root)
    # row 0 keeps the root's pose and velocity
    assert np.allclose(ww[:, 0], root[:, 10:13], atol=1e-12) and np.allclose(vc[:, 0], root[:, 7:10], atol=1e-12)
    gq = got[:, :L, 3:7] / np.linalg.norm(got[:, :L, 3:7], axis=-1, keepdims=True)
    assert np.abs(got[:, :L, :3] - pw).max() <= 4 * POSE_TOL_M
    assert np.abs(quat_to_R(gq) - Rw).max() <= 4 * POSE_TOL_RAD
    scale = np.abs(np.concatenate([vc, ww], -1)).reshape(B, -1).max(1)
    verr = np.abs(np.concatenate([got[:, :L, 7:10] - vc, got[:, :L, 10:13] - ww], -1)).reshape(B, -1).max(1)
    print(f"{tree}: link velocity rel err {(verr / scale).max():.2e}")
    assert (verr / scale).max() <= REL_TOL
    # poses only: the velocity slots keep their contents
    body2 = torch.full((B, max_links, 13), -5.0, device=DEV)
    body2[:, 0] = _t(root)
    m.compute_link_state_into(state[0, :, :n], None, body2)
    assert torch.equal(body2[:, :L, :7], body[:, :L, :7])
    assert bool((body2[:, 1:L, 7:] == -5.0).all()) and torch.equal(body2[:, 0, 7:], _t(root)[:, 7:])


def test_batched_calls_on_a_given_stream():
    """With stream= given, a batched call converts its inputs and allocates its
    outputs on that stream: while the current stream is held busy, FP64 device
    inputs (which need a conversion) give the same bits as the plain call."""
    rob = S.chain7()
    m, o = _pair(rob)
    q, qd, qdd, _ = S.samples(rob, 4096, 2300)
    ref = m.compute_fk_jacobian_rnea_batched(_t(q), _t(qd), _t(qdd))
    tgt = _t(o.link_poses(q)[:, 7])
    ref_ik = m.compute_inverse_kinematics_batched(7, tgt, _t(q + 0.05))
    q64, qd64, qdd64 = (torch.tensor(x, dtype=torch.float64, device=DEV) for x in (q, qd, qdd))
    q064 = torch.tensor(q + 0.05, dtype=torch.float64, device=DEV)
    tgt64 = tgt.double()
    torch.cuda.synchronize()  # the caller's inputs are ready before the current stream is held
    side = torch.cuda.Stream()
    torch.cuda._sleep(20_000_000)  # a bounded spin keeping the current stream busy (milliseconds)
    out = m.compute_fk_jacobian_rnea_batched(q64, qd64, qdd64, stream=side.cuda_stream)
    out_ik = m.compute_inverse_kinematics_batched(7, tgt64, q064, stream=side.cuda_stream)
    side.synchronize()
    torch.cuda.synchronize()
    for a, b in zip(out, ref):
        assert torch.equal(a, b)
    for a, b in zip(out_ik, ref_ik):
        assert torch.equal(a, b)


def test_multistart_ik():
    """Multi-start batched IK (mplib's Planner.IK(target, init_qpos,
    n_init_qpos)): the answer is the lowest start that converges, identical to
    running the starts one after the other with the single-start batched
    CLIK; targets no start reaches keep the smallest-error start."""
    import sapien_amd as sa

    rob = S.arm9()
    m, o = _pair(rob)
    link, B, n_init = rob.ee_link, 256, 6
    qs = S.samples(rob, B, 1960)[0]
    t = o.link_poses(qs)[:, link]
    t[::17, :3] += 5.0  # a few unreachable targets
    tt = _t(t)
    Q, ok, err, start = m.compute_inverse_kinematics_multistart_batched(link, tt, n_init_qpos=n_init,
                                                                        max_iterations=300, seed=3)
    # the same starts, one launch per start
    gen = torch.Generator(device=DEV)
    gen.manual_seed(3)
    lo, hi = (torch.as_tensor(x, dtype=torch.float32, device=DEV) for x in m.joint_limits())
    starts = torch.empty(1 + n_init, B, rob.nv, device=DEV)
    starts[0] = 0.0
    starts[1:] = lo + (hi - lo) * torch.rand(n_init, B, rob.nv, generator=gen, device=DEV)
    runs = [m.compute_inverse_kinematics_batched(link, tt, starts[k], max_iterations=300) for k in range(1 + n_init)]
    oks = torch.stack([r[1] for r in runs])
    first = torch.where(oks.any(0), oks.float().argmax(0), torch.stack([r[2].norm(dim=1) for r in runs]).argmin(0))
    assert torch.equal(start, first)
    for b in range(B):
        k = int(start[b])
        assert torch.equal(Q[b], runs[k][0][b]) and bool(ok[b]) == bool(runs[k][1][b])
        assert torch.equal(err[b], runs[k][2][b])
    unreachable = np.zeros(B, bool)
    unreachable[::17] = True
    assert not ok.cpu().numpy()[unreachable].any()
    assert ok.cpu().numpy()[~unreachable].mean() > 0.9
    assert oks[0].float().mean() <= ok.float().mean()  # extra starts never hurt
    # against the FP64 oracle: every converged answer reaches its target in the
    # oracle's FK, and on a 32-target sample the starts run in order through
    # oracle.clik pick the same start with the same outcome
    Qn, okn, startn = Q.cpu().numpy(), ok.cpu().numpy(), start.cpu().numpy()
    reached = o.link_poses(Qn[okn].astype(np.float64))[:, link]
    assert np.abs(reached[:, :3] - t[okn, :3]).max() < 2e-4
    dots = np.abs((reached[:, 3:] * t[okn, 3:]).sum(1)).clip(0, 1)
    assert (2 * np.arccos(dots)).max() < 2e-4
    # (an FP32 and an FP64 run from the same start part ways on slow problems,
    # so a differing choice must come from a start the GPU needed at least half
    # the iteration limit for, or one the oracle left within 10 eps)
    st = starts.cpu().numpy().astype(np.float64)
    tt32 = tt.cpu().numpy().astype(np.float64)
    its = torch.stack([r[3] for r in runs]).cpu().numpy()
    agree, slow = 0, []
    for b in np.random.default_rng(1961).choice(B, 32, replace=False):
        chosen, errs = None, []
        for k in range(1 + n_init):
            qr, okr, errr, _ = o.clik(link, tt32[b], q0=st[k, b], max_iterations=300)
            errs.append(np.linalg.norm(errr))
            if okr:
                chosen = k
                break
        gpu = int(startn[b]) if okn[b] else None
        if chosen == gpu:
            agree += 1
            continue
        k = min(x for x in (chosen, gpu) if x is not None)
        slow.append((int(b), gpu, chosen, int(its[k, b]), errs[k] if k < len(errs) else None))
        assert its[k, b] >= 150 or (k < len(errs) and errs[k] < 1e-3), slow[-1]
    print(f"multi-start: {agree}/32 same start as the oracle; differing (target, gpu, oracle, gpu iterations, "
          f"oracle error): {slow}")
    assert agree >= 28


def _robot_file(fname):
    import xml.etree.ElementTree as ET

    text = open(os.path.join(os.path.dirname(__file__), "golden", "robots", fname)).read()
    root = ET.fromstring(text)
    jn = [j.get("name") for j in root.findall("joint") if j.get("type") != "fixed"]
    ln = [l.get("name") for l in root.findall("link")]
    return text, jn, ln


ROBOT_FILES = sorted(f for f in os.listdir(os.path.join(os.path.dirname(__file__), "golden", "robots"))
                     if f.endswith(".urdf"))


@pytest.mark.parametrize("fname", ROBOT_FILES)
def test_real_robot_parity(fname):
    """SAPIEN's robot assets (fixed-joint chains, rpy inertial origins,
    continuous wheels, prismatic joints, massless links): the drop-in FP64
    per-call methods and the batched FP32 kernels against the oracle."""
    import sapien_amd as sa

    text, jn, ln = _robot_file(fname)
    m = sa.PinocchioModel(text, [0.0, 0.0, -9.81])
    m.set_joint_order(jn)
    m.set_link_order(ln)
    o = OracleModel(text)
    o.set_joint_order(jn)
    o.set_link_order(ln)
    n, B = len(jn), 512
    rng = np.random.default_rng(31)
    lo, hi = m.joint_limits()
    q = (lo + (hi - lo) * rng.uniform(size=(B, n))).astype(np.float32).astype(np.float64)
    qd = rng.normal(size=(B, n)).astype(np.float32).astype(np.float64)
    qdd = rng.normal(size=(B, n)).astype(np.float32).astype(np.float64)
    massive = np.abs(o.crba(q[:1])).max() > 0
    # forward dynamics only where M is positive definite (lwr and the gripper
    # have joints that move no mass: M is singular and ABA divides by zero)
    spd = massive and np.linalg.eigvalsh(o.crba(q[:16])).min() > 1e-12
    # drop-in FP64, one configuration per call
    for b in range(3):
        m.compute_forward_kinematics(q[b])
        P = np.array([np.concatenate([m.get_link_pose(i).p, m.get_link_pose(i).q]) for i in range(len(ln))])
        pos, ang = pose_errors(P[None], o.link_poses(q[b:b + 1]))
        assert pos <= 1e-6 and ang <= 1e-6
        assert np.allclose(m.compute_inverse_dynamics(q[b], qd[b], qdd[b]), o.rnea(q[b:b + 1], qd[b:b + 1],
                                                                                    qdd[b:b + 1])[0], atol=1e-9)
        m.compute_full_jacobian(q[b])
        for i in range(len(ln)):
            assert np.allclose(m.get_link_jacobian(i), o.link_jacobian(q[b], i)[0], atol=1e-9)
    # batched FP32, both kernel families
    for v in ("auto", "thread_per_env"):
        sa.set_kernel_variant(m, v)
        _check_kin(m, o, q, qd, qdd)
        if massive:
            assert rel_err(m.compute_generalized_mass_matrix_batched(_t(q)).cpu().numpy(), o.crba(q)) <= REL_TOL
        if spd:
            tau = o.rnea(q, qd, qdd)
            A = m.compute_forward_dynamics_batched(_t(q), _t(qd), _t(tau)).cpu().numpy()
            assert rel_err(A, qdd) <= 1e-3, (v, sa.dynamics_precision(m))


def test_clik_mixed_workload_size():
    """The benchmark's CLIK workload at its size: 65,536 problems of the four
    trees (4:3:2:1), targets = FK of seeded configurations, q0 = neutral. The
    mixed launch equals each model's own batched CLIK bit for bit, every
    problem reports at most max_iterations, a converged problem's error is
    below eps, and a sample of converged results reaches its target in the
    FP64 oracle."""
    import sapien_amd as sa

    counts = [26214, 19661, 13107, 6554]
    models, links, tgts = [], [], []
    for i, (tree, nprob) in enumerate(zip(["chain7", "arm9", "quad18", "torso30"], counts)):
        rob = S.robot(tree)
        m, o = _pair(rob)
        q = _t(S.samples(rob, nprob, 1004 + i)[0])
        models.append((m, o, rob))
        links.append(rob.ee_link)
        tgts.append(m.compute_forward_kinematics_batched(q)[:, rob.ee_link].contiguous())
    outs = sa.clik_hetero_batched([x[0] for x in models], links, tgts)
    total_ok = 0
    for (m, o, rob), link, t, (Q, ok, err, it) in zip(models, links, tgts, outs):
        Qr, okr, errr, itr = m.compute_inverse_kinematics_batched(link, t)
        assert torch.equal(Q, Qr) and torch.equal(ok, okr) and torch.equal(err, errr) and torch.equal(it, itr), rob.name
        okn, itn, errn = ok.cpu().numpy(), it.cpu().numpy(), err.cpu().numpy()
        assert ((itn >= 0) & (itn <= 1000)).all() and (itn[~okn] == 1000).all()
        assert (np.linalg.norm(errn[okn], axis=1) < 1e-4).all()
        idx = np.flatnonzero(okn)[:256]
        reached = o.link_poses(Q.cpu().numpy()[idx])[:, link, :3]
        assert np.abs(reached - t.cpu().numpy()[idx, :3]).max() < 2e-4, rob.name
        total_ok += okn.sum()
    assert total_ok / sum(counts) > 0.9


@pytest.mark.parametrize("tree", ["chain7", "torso30"])
def test_random_qpos(tree):
    """get_random_qpos_batched: the numpy Philox4x32-10 restatement's values
    (FP32-rounded, bit for bit), the same batch for the same seed, every
    entry within the joint limits (torso30's continuous wheels in [-pi, pi)),
    moments of a uniform distribution; get_random_qpos (per call, FP64 on the
    GPU) within the limits. The reference's own values come from Pinocchio's
    generator and are not reproducible without it."""
    from philox import random_qpos

    rob = S.robot(tree)
    m, _ = _pair(rob)
    lo, hi = m.joint_limits()
    B = 20000
    a = m.get_random_qpos_batched(B, seed=12345).cpu().numpy()
    ref = random_qpos(B, lo, hi, 12345).astype(np.float32)
    assert np.array_equal(a, ref)
    assert torch.equal(m.get_random_qpos_batched(B, seed=12345), m.get_random_qpos_batched(B, seed=12345))
    assert not np.array_equal(m.get_random_qpos_batched(16, seed=7).cpu().numpy(), a[:16])
    assert ((a >= lo.astype(np.float32)) & (a <= hi.astype(np.float32))).all()
    u = (a - lo) / (hi - lo)  # uniform on [0, 1): mean 1/2, variance 1/12
    assert np.abs(u.mean(0) - 0.5).max() < 0.01 and np.abs(u.var(0) - 1 / 12).max() < 0.005
    for _ in range(5):
        q = m.get_random_qpos()
        assert q.dtype == np.float64 and q.shape == (rob.nv,) and ((q >= lo) & (q <= hi)).all()


def test_tree_groups_and_clik_schedules():
    """Every autotuner choice computes the same thing: the mixed-tree launch
    with every lanes-per-environment choice of each segment against the FP64
    oracle (FP32 tolerances), and batched CLIK under every schedule: converged
    problems reach their targets, the success rate stays, and a mixed-tree
    launch equals the per-model launches bit for bit."""
    import sapien_amd as sa
    from sapien_amd import tune

    trees = ["chain7", "arm9", "quad18", "torso30"]
    pairs = [_pair(S.robot(t)) for t in trees]
    Bs = [300, 200, 100, 50]
    ins = [S.samples(S.robot(t), b, 2400 + i) for i, (t, b) in enumerate(zip(trees, Bs))]
    try:
        for G7 in (8, 16, 32):
            for G9 in (16, 32):
                tune.set_tree_group(pairs[0][0], G7)
                tune.set_tree_group(pairs[1][0], G9)
                outs = sa.hetero_batched([p[0] for p in pairs], [_t(x[0]) for x in ins], [_t(x[1]) for x in ins],
                                         [_t(x[2]) for x in ins])
                for (m, o), (q, qd, qdd, _), d in zip(pairs, ins, outs):
                    pos, ang = pose_errors(d["poses"].cpu().numpy(), o.link_poses(q))
                    assert pos <= POSE_TOL_M and ang <= POSE_TOL_RAD, (G7, G9)
                    assert rel_err(d["J"].cpu().numpy(), o.jacobian_world(q)) <= REL_TOL
                    assert rel_err(d["tau"].cpu().numpy(), o.rnea(q, qd, qdd)) <= REL_TOL
                    assert rel_err(d["M"].cpu().numpy(), o.crba(q)) <= REL_TOL
        with pytest.raises(ValueError):
            tune.set_tree_group(pairs[1][0], 8)  # fewer lanes than joints
    finally:
        for m, _ in pairs:
            tune.set_tree_group(m, 0)
    # CLIK schedules
    links = [S.robot(t).ee_link for t in trees]
    tg = [_t(o.link_poses(S.samples(S.robot(t), n, 2410 + i)[0])[:, l])
          for i, ((m, o), t, n, l) in enumerate(zip(pairs, trees, [400, 300, 200, 100], links))]
    try:
        for sch in tune.CLIK_SCHEDULES:
            tune.set_clik_schedule(*sch)
            assert tune.get_clik_schedule() == sch
            mixed = sa.clik_hetero_batched([p[0] for p in pairs], links, tg)
            for (m, o), l, t, (Q, ok, err, it) in zip(pairs, links, tg, mixed):
                Qr, okr, errr, itr = m.compute_inverse_kinematics_batched(l, t)
                assert torch.equal(Q, Qr) and torch.equal(ok, okr) and torch.equal(it, itr), sch
                okn = ok.cpu().numpy().astype(bool)
                assert okn.mean() > 0.5, sch
                reached = o.link_poses(Q.cpu().numpy()[okn].astype(np.float64))[:, l, :3]
                assert np.abs(reached - t.cpu().numpy()[okn, :3]).max() < 2e-4, sch
        with pytest.raises(ValueError):
            tune.set_clik_schedule(128, 5, 8)
    finally:
        tune.set_clik_schedule(128, 4, 8)
