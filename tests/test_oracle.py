"""Pins the FP64 oracle before anything is compared against it.

Every oracle function is checked against something it does not share code with:
the reference's own known-answer tests (unittest/test_sapien/test_math.py),
brute-force Lagrangian mechanics of the raw URDF links (tests/bruteforce.py),
closed forms (pendulum, Rodrigues), exact dual-number derivatives versus
finite differences, and algebraic invariants chosen so that a dropped term,
a wrong sign or a transposed operand fails.
"""
import math
import os

import numpy as np
import pytest

from bruteforce import BruteRobot
from exporter_style import rewrite
from oracle.urdf_model import OracleModel, log6, log6_mat, rpy_to_matrix
from sapien_amd import synthetic as S

GOLD = os.path.join(os.path.dirname(__file__), "golden")
TREES = list(S.ROBOTS)


def _model(rob, urdf=None):
    m = OracleModel(urdf or rob.urdf)
    m.set_joint_order(rob.joint_names)
    m.set_link_order(rob.link_names)
    return m


def _quat_to_R(q):
    w, x, y, z = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                     [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                     [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])


# ----------------------------------------------------------------- reference KATs (test_math.py)
def test_kat_rpy_to_quat():
    # unittest/test_sapien/test_math.py:58-69: rpy [0.3, 0.4, 0.5] -> quaternion (wxyz)
    urdf = ('<robot name="r"><link name="a"/><link name="b"/><joint name="j" type="fixed">'
            '<origin xyz="0 0 0" rpy="0.3 0.4 0.5"/><parent link="a"/><child link="b"/></joint>'
            '<link name="c"/><joint name="k" type="revolute"><parent link="b"/><child link="c"/>'
            '<limit lower="-1" upper="1"/></joint></robot>')
    m = OracleModel(urdf)
    m.set_joint_order(["k"])
    m.set_link_order(["b"])
    q = m.link_poses(np.zeros(1))[0, 0, 3:]
    ref = np.array([0.94628083, 0.09330659, 0.22656631, 0.21098383])
    assert np.allclose(q, ref, atol=1e-7) or np.allclose(q, -ref, atol=1e-7)


def test_kat_matrix_to_quat_and_back():
    # unittest/test_sapien/test_math.py:27-42 and :79-95
    R = np.array([[-0.1301451, -0.5405447, 0.8311881], [-0.9910400, 0.0963137, -0.0925388],
                  [-0.0300335, -0.8357841, -0.5482362]])
    ref = np.array([-0.3232385, 0.5748429, -0.6660884, 0.3484233])
    assert np.allclose(_quat_to_R(ref), R, atol=1e-6)
    # quaternion extraction through a fixed joint whose rpy encodes R
    p = math.asin(-R[2, 0]); r = math.atan2(R[2, 1], R[2, 2]); y = math.atan2(R[1, 0], R[0, 0])
    assert np.allclose(rpy_to_matrix(r, p, y), R, atol=1e-6)
    urdf = (f'<robot name="r"><link name="a"/><link name="b"/><joint name="j" type="fixed">'
            f'<origin xyz="0.1 0.2 0.3" rpy="{r} {p} {y}"/><parent link="a"/><child link="b"/></joint>'
            f'<link name="c"/><joint name="k" type="prismatic"><parent link="b"/><child link="c"/></joint></robot>')
    m = OracleModel(urdf)
    m.set_joint_order(["k"])
    m.set_link_order(["b"])
    pose = m.link_poses(np.zeros(1))[0, 0]
    assert np.allclose(pose[:3], [0.1, 0.2, 0.3])
    assert np.allclose(pose[3:], ref, atol=1e-6) or np.allclose(pose[3:], -ref, atol=1e-6)


def test_quat_extraction_all_branches():
    """Shoemake: every branch (trace>0, and max diagonal x/y/z) reproduces R."""
    rng = np.random.default_rng(7)
    for angle, axis in [(0.3, [0, 0, 1]), (3.0, [1, 0, 0]), (3.0, [0, 1, 0]), (3.0, [0, 0, 1]), (math.pi, [1, 1, 0])]:
        a = np.array(axis, float) / np.linalg.norm(axis)
        K = np.array([[0, -a[2], a[1]], [a[2], 0, -a[0]], [-a[1], a[0], 0]])
        R = np.eye(3) + math.sin(angle) * K + (1 - math.cos(angle)) * K @ K
        r = math.atan2(R[2, 1], R[2, 2]); p = math.asin(max(-1, min(1, -R[2, 0]))); y = math.atan2(R[1, 0], R[0, 0])
        if not np.allclose(rpy_to_matrix(r, p, y), R, atol=1e-9):
            continue
        urdf = (f'<robot name="r"><link name="a"/><link name="b"/><joint name="j" type="fixed">'
                f'<origin rpy="{r!r} {p!r} {y!r}"/><parent link="a"/><child link="b"/></joint>'
                f'<link name="c"/><joint name="k" type="prismatic"><parent link="b"/><child link="c"/></joint></robot>')
        m = OracleModel(urdf)
        m.set_joint_order(["k"])
        m.set_link_order(["b"])
        qq = m.link_poses(np.zeros(1))[0, 0, 3:]
        assert abs(np.linalg.norm(qq) - 1) < 1e-12
        assert np.allclose(_quat_to_R(qq), R, atol=1e-9)


# ----------------------------------------------------------------- FK against brute force
def _all_cases():
    cases = [(n, S.robot(n), None) for n in TREES]
    cases += [(n + "-exported", S.robot(n), rewrite(S.robot(n).urdf, seed=5)) for n in ("chain7", "torso30")]
    return cases


@pytest.mark.parametrize("name,rob,urdf", _all_cases(), ids=[c[0] for c in _all_cases()])
def test_fk_matches_bruteforce(name, rob, urdf):
    m = _model(rob, urdf)
    br = BruteRobot(urdf or rob.urdf, rob.joint_names)
    q = S.samples(rob, 4, 11)[0]
    R, p = m.link_mats(q)
    for b in range(4):
        T, _ = br.transforms(q[b])
        for l, ln in enumerate(rob.link_names):
            assert np.allclose(T[ln][:3, 3], p[b, l], atol=1e-12)
            assert np.allclose(T[ln][:3, :3], R[b, l], atol=1e-12)


def test_fk_movo_fixture_and_panda():
    for fname in ("movo_simple.urdf", "panda.urdf"):
        text = open(os.path.join(GOLD, fname)).read()
        import xml.etree.ElementTree as ET
        root = ET.fromstring(text)
        jn = [j.get("name") for j in root.findall("joint") if j.get("type") != "fixed"]
        ln = [l.get("name") for l in root.findall("link")]
        m = OracleModel(text)
        m.set_joint_order(jn)
        m.set_link_order(ln)
        br = BruteRobot(text, jn)
        # test_articulation.py:80 qpos for movo
        q = np.array([0.1, 0.1, 0.1, 0.2, 0.2, 0.2, 0.2, 0.2, 0.2, 0.2, 0.3, 0.4, 0.5]) if len(jn) == 13 else \
            np.random.default_rng(3).uniform(-1, 1, len(jn))
        R, p = m.link_mats(q)
        T, _ = br.transforms(q)
        for l, name in enumerate(ln):
            assert np.allclose(T[name][:3, 3], p[0, l], atol=1e-12), name
            assert np.allclose(T[name][:3, :3], R[0, l], atol=1e-12), name


def test_fk_zero_config_is_sum_of_offsets():
    rob = S.chain7()
    m = _model(rob)
    import xml.etree.ElementTree as ET
    offs = [np.array([float(x) for x in j.find("origin").get("xyz").split()]) for j in ET.fromstring(rob.urdf).findall("joint")]
    P = m.link_poses(np.zeros(7))[0]
    assert np.allclose(P[1:, :3], np.cumsum(offs, axis=0), atol=1e-14)
    assert np.allclose(np.abs(P[:, 3]), 1.0)


def test_fk_single_revolute_rodrigues():
    urdf = ('<robot name="r"><link name="a"/><link name="b"/><link name="c"/>'
            '<joint name="j" type="revolute"><origin xyz="0 0 0"/><axis xyz="1 2 3"/><parent link="a"/><child link="b"/>'
            '<limit lower="-3" upper="3"/></joint><joint name="f" type="fixed"><origin xyz="0.5 -0.2 0.1"/>'
            '<parent link="b"/><child link="c"/></joint></robot>')
    m = OracleModel(urdf)
    m.set_joint_order(["j"])
    m.set_link_order(["c"])
    t = 0.77
    a = np.array([1, 2, 3.0]) / math.sqrt(14)
    # Rodrigues via quaternion: q = (cos t/2, a sin t/2)
    qr = np.r_[math.cos(t / 2), a * math.sin(t / 2)]
    R = _quat_to_R(qr)
    P = m.link_poses(np.array([t]))[0, 0]
    assert np.allclose(P[:3], R @ [0.5, -0.2, 0.1], atol=1e-14)
    assert np.allclose(P[3:], qr, atol=1e-14) or np.allclose(P[3:], -qr, atol=1e-14)


# ----------------------------------------------------------------- Jacobians
@pytest.mark.parametrize("tree", TREES)
def test_jacobian_matches_exact_fk_derivative(tree):
    rob = S.robot(tree)
    m = _model(rob)
    q = S.samples(rob, 3, 12)[0]
    R, p = m.link_mats(q)
    for l in range(rob.nlinks):
        dp = m.link_pos_jac(q, l)  # exact d p_link / dq
        Jw = m.link_jacobian(q, l, local=False)
        Jl = m.link_jacobian(q, l, local=True)
        for b in range(3):
            # world Jacobian: linear rows = velocity of the world-origin point (appendix A.4)
            pred = Jw[b, :3] - np.cross(p[b, l], Jw[b, 3:].T).T
            assert np.allclose(pred, dp[b], atol=1e-12)
            # local = body frame: R^T * (velocity of the link origin), R^T * omega
            assert np.allclose(Jl[b, :3], R[b, l].T @ dp[b], atol=1e-12)
            assert np.allclose(Jl[b, 3:], R[b, l].T @ Jw[b, 3:], atol=1e-12)


def test_jacobian_angular_rows_by_finite_differences():
    rob = S.arm9()
    m = _model(rob)
    q = S.samples(rob, 1, 13)[0][0]
    h = 1e-6
    J = m.link_jacobian(q, 8, local=False)[0]
    R0 = m.link_mats(q)[0][0, 8]
    for k in range(rob.nv):
        e = np.zeros(rob.nv); e[k] = h
        dR = (m.link_mats(q + e)[0][0, 8] - m.link_mats(q - e)[0][0, 8]) / (2 * h)
        W = dR @ R0.T  # skew(omega)
        w = np.array([W[2, 1], W[0, 2], W[1, 0]])
        assert np.allclose(w, J[3:, k], atol=1e-8)


def test_full_jacobian_columns_and_support():
    rob = S.torso30()
    m = _model(rob)
    q = S.samples(rob, 2, 14)[0]
    Jfull = m.jacobian_world(q)
    for l in range(rob.nlinks):
        Jl = m.link_jacobian(q, l)
        nz = np.abs(Jl).sum(axis=1) > 0
        # a link Jacobian is a column-masked view of the full Jacobian
        assert np.allclose(Jl[nz[:, None, :].repeat(6, 1)], Jfull[nz[:, None, :].repeat(6, 1)])
    # the root link has no support
    assert np.all(m.link_jacobian(q, 0) == 0)


# ----------------------------------------------------------------- dynamics against brute force
@pytest.mark.parametrize("tree", TREES)
def test_crba_matches_bruteforce_mass_matrix(tree):
    rob = S.robot(tree)
    m = _model(rob)
    br = BruteRobot(rob.urdf, rob.joint_names)
    q = S.samples(rob, 2, 15)[0]
    M = m.crba(q)
    for b in range(2):
        assert np.allclose(M[b], br.mass_matrix(q[b]), atol=1e-12, rtol=1e-10)


@pytest.mark.parametrize("tree", TREES)
def test_rnea_gravity_is_potential_gradient(tree):
    rob = S.robot(tree)
    m = _model(rob)
    br = BruteRobot(rob.urdf, rob.joint_names)
    q = S.samples(rob, 2, 16)[0]
    z = np.zeros_like(q)
    tau = m.rnea(q, z, z)
    h = 1e-6
    for b in range(2):
        assert np.allclose(tau[b], br.gravity_torque(q[b]), atol=1e-11)
        fd = np.array([(br.potential(q[b] + h * e) - br.potential(q[b] - h * e)) / (2 * h) for e in np.eye(rob.nv)])
        assert np.allclose(tau[b], fd, atol=1e-7)


@pytest.mark.parametrize("tree", ["chain7", "arm9", "quad18"])
def test_rnea_matches_lagrangian(tree):
    rob = S.robot(tree)
    m = _model(rob)
    br = BruteRobot(rob.urdf, rob.joint_names)
    q, qd, qdd, _ = S.samples(rob, 2, 17)
    tau = m.rnea(q, qd, qdd)
    for b in range(2):
        ref = br.inverse_dynamics(q[b], qd[b], qdd[b])
        assert np.allclose(tau[b], ref, atol=1e-6, rtol=1e-7)


def test_exported_layout_gives_identical_dynamics():
    """The exporter's layout (dummy links, x-axis joint frames, principal-axis
    inertials) is the same mechanism: identical outputs to rounding when the
    values are printed in full, and within the FP32 quantization of the
    geometry (about 6e-8 relative per value) as the exporter prints them."""
    from parity import rel_err

    for tree in ("torso30", "arm9"):
        rob = S.robot(tree)
        m1 = _model(rob)
        q, qd, qdd, tau = S.samples(rob, 3, 18)
        m2 = _model(rob, rewrite(rob.urdf, seed=9, fp32=False))
        assert np.allclose(m1.rnea(q, qd, qdd), m2.rnea(q, qd, qdd), atol=1e-11)
        assert np.allclose(m1.crba(q), m2.crba(q), atol=1e-12)
        assert np.allclose(m1.jacobian_world(q), m2.jacobian_world(q), atol=1e-12)
        assert np.allclose(m1.link_poses(q)[..., :3], m2.link_poses(q)[..., :3], atol=1e-12)
        m3 = _model(rob, rewrite(rob.urdf, seed=9))
        assert rel_err(m3.rnea(q, qd, qdd), m1.rnea(q, qd, qdd)) < 1e-5
        assert rel_err(m3.crba(q), m1.crba(q)) < 1e-5
        assert rel_err(m3.jacobian_world(q), m1.jacobian_world(q)) < 1e-5
        assert np.abs(m3.link_poses(q)[..., :3] - m1.link_poses(q)[..., :3]).max() < 1e-6


def test_pendulum_closed_form():
    m_, l_, Izz = 1.7, 0.6, 0.013
    urdf = (f'<robot name="p"><link name="base"/><link name="bob"><inertial><origin xyz="{l_} 0 0"/>'
            f'<mass value="{m_}"/><inertia ixx="0.01" iyy="0.02" izz="{Izz}" ixy="0" ixz="0" iyz="0"/></inertial></link>'
            f'<joint name="j" type="revolute"><axis xyz="0 0 1"/><parent link="base"/><child link="bob"/>'
            f'<limit lower="-3" upper="3"/></joint></robot>')
    m = OracleModel(urdf, gravity=(0, -9.81, 0))
    m.set_joint_order(["j"])
    m.set_link_order(["bob"])
    for q, qd, qdd in [(0.3, 0.0, 0.0), (1.1, -2.0, 0.5), (-2.5, 1.0, -3.0)]:
        tau = m.rnea([q], [qd], [qdd])[0, 0]
        ref = (Izz + m_ * l_ ** 2) * qdd + m_ * 9.81 * l_ * math.cos(q)
        assert abs(tau - ref) < 1e-12
        assert abs(m.crba([q])[0, 0, 0] - (Izz + m_ * l_ ** 2)) < 1e-14
        assert abs(m.aba([q], [qd], [ref])[0, 0] - qdd) < 1e-12


@pytest.mark.parametrize("tree", TREES)
def test_crba_columns_are_rnea_differences_and_spd(tree):
    rob = S.robot(tree)
    m = _model(rob)
    q = S.samples(rob, 1, 19)[0]
    n = rob.nv
    z = np.zeros((n, n))
    Q = np.repeat(q, n, 0)
    cols = m.rnea(Q, z, np.eye(n)) - m.rnea(Q, z, z)
    M = m.crba(q)[0]
    assert np.allclose(cols.T, M, atol=1e-11)
    assert np.allclose(M, M.T)
    assert np.linalg.eigvalsh(M).min() > 0


@pytest.mark.parametrize("tree", TREES)
def test_aba_inverts_rnea_and_matches_dense_solve(tree):
    rob = S.robot(tree)
    m = _model(rob)
    q, qd, qdd, tau = S.samples(rob, 4, 20)
    assert np.allclose(m.aba(q, qd, m.rnea(q, qd, qdd)), qdd, atol=1e-9)
    M = m.crba(q)
    b = m.rnea(q, qd, np.zeros_like(q))
    dense = np.linalg.solve(M, (tau - b)[..., None])[..., 0]
    assert np.allclose(m.aba(q, qd, tau), dense, atol=1e-8, rtol=1e-8)


@pytest.mark.parametrize("tree", ["chain7", "arm9", "torso30"])
def test_coriolis_properties(tree):
    rob = S.robot(tree)
    m = _model(rob)
    q, qd, _, _ = S.samples(rob, 2, 21)
    C = m.coriolis(q, qd)
    z = np.zeros_like(q)
    bq = m.rnea(q, qd, z) - m.rnea(q, z, z)
    assert np.allclose(np.einsum("bij,bj->bi", C, qd), bq, atol=1e-10)
    h = 1e-6
    for b in range(2):
        Mdot = (m.crba(q[b] + h * qd[b])[0] - m.crba(q[b] - h * qd[b])[0]) / (2 * h)
        assert np.allclose(Mdot, C[b] + C[b].T, atol=1e-7)
        N = Mdot - 2 * C[b]
        assert np.allclose(N, -N.T, atol=1e-7)


@pytest.mark.parametrize("tree", ["chain7", "quad18"])
def test_rnea_derivatives_match_finite_differences(tree):
    rob = S.robot(tree)
    m = _model(rob)
    q, qd, qdd, _ = S.samples(rob, 1, 22)
    dq, dqd, dqdd = m.rnea_derivs(q, qd, qdd)
    h = 1e-6
    n = rob.nv
    for k in range(n):
        e = np.zeros((1, n)); e[0, k] = h
        assert np.allclose((m.rnea(q + e, qd, qdd) - m.rnea(q - e, qd, qdd))[0] / (2 * h), dq[0, :, k], atol=1e-6)
        assert np.allclose((m.rnea(q, qd + e, qdd) - m.rnea(q, qd - e, qdd))[0] / (2 * h), dqd[0, :, k], atol=1e-6)
    assert np.allclose(dqdd[0], m.crba(q)[0], atol=1e-12)


# ----------------------------------------------------------------- SE(3) log and CLIK
def _exp6(xi):
    v, w = xi[:3], xi[3:]
    t = np.linalg.norm(w)
    K = np.array([[0, -w[2], w[1]], [w[2], 0, -w[0]], [-w[1], w[0], 0]])
    if t < 1e-12:
        return np.eye(3) + K, v
    R = np.eye(3) + math.sin(t) / t * K + (1 - math.cos(t)) / t ** 2 * K @ K
    V = np.eye(3) + (1 - math.cos(t)) / t ** 2 * K + (t - math.sin(t)) / t ** 3 * K @ K
    return R, V @ v


def test_log6_inverts_exp6_and_jlog6_is_its_derivative():
    rng = np.random.default_rng(23)
    for t in [1e-7, 1e-3, 0.5, 2.0, 3.0, 3.13]:
        w = rng.normal(size=3)
        w *= t / np.linalg.norm(w)
        xi = np.r_[rng.normal(size=3), w]
        R, p = _exp6(xi)
        l, J = log6_mat(R, p)
        assert np.allclose(l, xi, atol=1e-8 if t > 3 else 1e-11), t
        # Jlog6 = d log6(M exp(d)) / dd at d = 0 (right perturbation)
        h = 1e-6
        for k in range(6):
            d = np.zeros(6); d[k] = h
            Rp, pp = _exp6(d)
            lp, _ = log6_mat(R @ Rp, R @ pp + p)
            Rm, pm = _exp6(-d)
            lm, _ = log6_mat(R @ Rm, R @ pm + p)
            assert np.allclose((lp - lm) / (2 * h), J[:, k], atol=2e-6 if t > 3 else 1e-7), (t, k)


def test_target_quaternion_is_not_normalized():
    """The CLIK target is built as Eigen::Quaterniond(w, x, y, z).toRotationMatrix()
    (pinocchio_model.cpp:255), which applies the unit-quaternion formula to q as
    given: q = s q_hat gives R = (1 - s^2) I + s^2 R(q_hat). For q_hat a rotation
    by phi about k that matrix has trace 3 - 2 s^2 (1 - cos phi) and skew part
    s^2 sin(phi) [k]x, so the trace-based log3 returns
    theta / sin(theta) * s^2 sin(phi) k with cos(theta) = 1 - s^2 (1 - cos phi).
    A normalizing conversion would return phi k instead."""
    k = np.array([0.3, -0.5, 0.81])
    k /= np.linalg.norm(k)
    for phi, s in [(0.7, 0.9975), (2.0, 1.02), (0.3, 1.0)]:
        qv = s * np.r_[math.cos(phi / 2), math.sin(phi / 2) * k]
        l, _ = log6(np.r_[0.0, 0.0, 0.0, qv], with_jac=False)
        theta = math.acos(1 - s * s * (1 - math.cos(phi)))
        ref = theta / math.sin(theta) * s * s * math.sin(phi) * k
        assert np.allclose(l[3:], ref, atol=1e-13), (phi, s)
        assert np.allclose(l[:3], 0, atol=1e-15)
    # the reference's own panda target (manualtest/urdf.py:32-33) has |q|^2 = 0.9951
    assert abs(np.linalg.norm([0.21, 0.73, 0.41, 0.5]) ** 2 - 0.9951) < 1e-12


@pytest.mark.parametrize("tree", ["chain7", "arm9"])
def test_clik_converges_to_fk_targets(tree):
    rob = S.robot(tree)
    m = _model(rob)
    q = S.samples(rob, 6, 24)[0]
    P = m.link_poses(q)
    ok_count = 0
    for b in range(6):
        target = P[b, rob.ee_link]
        qo, ok, err, it = m.clik(rob.ee_link, target, q0=q[b] + 0.2)
        if ok:
            ok_count += 1
            assert np.linalg.norm(err) < 1e-4
            Pr = m.link_poses(qo)[0, rob.ee_link]
            assert np.allclose(Pr[:3], target[:3], atol=1e-4)
    assert ok_count >= 4


def test_clik_solved_target_mask_and_unreachable():
    rob = S.chain7()
    m = _model(rob)
    q = S.samples(rob, 1, 25)[0][0]
    target = m.link_poses(q)[0, 7]
    qo, ok, err, it = m.clik(7, target, q0=q)
    assert ok and it == 0 and np.allclose(qo, q)
    mask = np.array([1, 1, 1, 0, 1, 1, 0])
    qo, ok, err, it = m.clik(7, m.link_poses(q + 0.1 * mask)[0, 7], q0=q, mask=mask)
    assert qo[3] == q[3] and qo[6] == q[6]
    far = target.copy()
    far[:3] = [10.0, 10.0, 10.0]
    qo, ok, err, it = m.clik(7, far, q0=q, max_iterations=50)
    assert not ok and it == 50 and np.linalg.norm(err) > 1


def test_rollout_jvp_matches_finite_differences():
    rob = S.chain7()
    m = _model(rob)
    rng = np.random.default_rng(26)
    q0, qd0, _, _ = S.samples(rob, 1, 27)
    T, dt = 16, 1.0 / 240
    tau = rng.normal(0, math.sqrt(5), size=(1, T, 7))
    dq0, dqd0, dtau = rng.normal(size=7), rng.normal(size=7), rng.normal(size=(T, 7))
    jq, jqd = m.rollout_jvp(q0[0], qd0[0], tau[0], dq0, dqd0, dtau, dt)
    h = 1e-6
    a = m.rollout(q0 + h * dq0, qd0 + h * dqd0, tau + h * dtau[None], dt)
    b = m.rollout(q0 - h * dq0, qd0 - h * dqd0, tau - h * dtau[None], dt)
    assert np.allclose((a[0] - b[0])[0] / (2 * h), jq, atol=1e-6)
    assert np.allclose((a[1] - b[1])[0] / (2 * h), jqd, atol=1e-6)


def _slider_urdf(mass, axis="1 0 0"):
    # one body on a prismatic joint; its axis is orthogonal to gravity (0, 0, -9.81)
    return ('<robot name="slider"><link name="base"/>'
            f'<link name="body"><inertial><origin xyz="0.1 0.2 0.3"/><mass value="{mass!r}"/>'
            '<inertia ixx="0.01" iyy="0.02" izz="0.03" ixy="0" ixz="0" iyz="0"/></inertial></link>'
            '<joint name="slide" type="prismatic"><parent link="base"/><child link="body"/>'
            f'<origin xyz="0.5 -0.2 1.0"/><axis xyz="{axis}"/><limit lower="-10" upper="10"/></joint></robot>')


def test_rollout_is_semi_implicit_euler_closed_form():
    """A 1-dof prismatic body of mass m under a constant force F along an axis
    orthogonal to gravity has qdd = F/m exactly, so T steps of semi-implicit
    Euler (qd += dt qdd; then q += dt qd) give
        qd_T = qd0 + T dt F/m,   q_T = q0 + T dt qd0 + dt^2 (F/m) T (T + 1) / 2.
    Explicit Euler (q updated with the old qd) would give T (T - 1) / 2, and a
    dropped or doubled dt changes both lines."""
    mass, F, T, dt = 1.7, 3.1, 37, 1.0 / 240
    for axis in ("1 0 0", "0 1 0", "0.6 0.8 0"):
        m = OracleModel(_slider_urdf(mass, axis))
        m.set_joint_order(["slide"])
        m.set_link_order(["body"])
        q0, qd0 = np.array([[0.25], [-0.4]]), np.array([[0.5], [-1.25]])
        tau = np.full((2, T, 1), F)
        qT, qdT = m.rollout(q0, qd0, tau, dt)
        a = F / mass
        assert np.allclose(qdT, qd0 + T * dt * a, rtol=0, atol=1e-13)
        assert np.allclose(qT, q0 + T * dt * qd0 + dt * dt * a * T * (T + 1) / 2, rtol=0, atol=1e-13)
        # qdd from the oracle's forward dynamics is exactly F/m (no gravity along the axis)
        assert np.allclose(m.aba(q0, qd0, tau[:, 0]), a, rtol=1e-14)


def test_rollout_equals_step_loop_over_forward_dynamics():
    """oracle.rollout is T applications of qd += dt ABA(q, qd, tau_t), q += dt qd
    over the oracle's ABA, which the ABA tests pin (RNEA round trip, dense solve)."""
    for tree in ("chain7", "quad18"):
        rob = S.robot(tree)
        m = _model(rob)
        q0, qd0, _, _ = S.samples(rob, 3, 31)
        T, dt = 9, 1.0 / 240
        tau = np.random.default_rng(32).normal(0, 5, size=(3, T, rob.nv))
        qT, qdT = m.rollout(q0, qd0, tau, dt)
        q, qd = q0.copy(), qd0.copy()
        for t in range(T):
            qd = qd + dt * m.aba(q, qd, tau[:, t])
            q = q + dt * qd
        assert np.allclose(qT, q, rtol=1e-12, atol=1e-12)
        assert np.allclose(qdT, qd, rtol=1e-12, atol=1e-12)


def test_error_semantics():
    rob = S.chain7()
    m = OracleModel(rob.urdf)
    with pytest.raises(ValueError):
        m.set_joint_order(["nope"])
    with pytest.raises(RuntimeError):
        m.set_joint_order(rob.joint_names[:-1])
    with pytest.raises(ValueError):
        m.set_link_order(["nope"])


# ----------------------------------------------------------------- real robots (URDF ingest)
ROBOT_FILES = sorted(f for f in os.listdir(os.path.join(GOLD, "robots")) if f.endswith(".urdf"))


def _robot_file(fname):
    import xml.etree.ElementTree as ET

    text = open(os.path.join(GOLD, "robots", fname)).read()
    root = ET.fromstring(text)
    jn = [j.get("name") for j in root.findall("joint") if j.get("type") != "fixed"]
    ln = [l.get("name") for l in root.findall("link")]
    return text, jn, ln


@pytest.mark.parametrize("fname", ROBOT_FILES)
def test_real_robots_match_bruteforce(fname):
    """SAPIEN's own robot assets through the oracle's URDF reader: link poses,
    mass matrix and gravity torques against the brute-force link-by-link model."""
    text, jn, ln = _robot_file(fname)
    m = OracleModel(text)
    m.set_joint_order(jn)
    m.set_link_order(ln)
    br = BruteRobot(text, jn)
    q = np.random.default_rng(21).uniform(-1, 1, (3, len(jn)))
    R, p = m.link_mats(q)
    z = np.zeros((1, len(jn)))
    for b in range(3):
        T, _ = br.transforms(q[b])
        for l, name in enumerate(ln):
            assert np.allclose(T[name][:3, 3], p[b, l], atol=1e-12), name
            assert np.allclose(T[name][:3, :3], R[b, l], atol=1e-12), name
        Mb = br.mass_matrix(q[b])
        assert np.allclose(m.crba(q[b:b + 1])[0], Mb, rtol=1e-10, atol=1e-12)
        assert np.allclose(m.rnea(q[b:b + 1], z, z)[0], br.gravity_torque(q[b]), rtol=1e-10, atol=1e-12)


@pytest.mark.parametrize("tree", ["arm9", "quad18"])
def test_link_state_reference_is_the_time_derivative_of_the_poses(tree):
    """The link-state reference the simulator-adapter test uses: its
    centre-of-mass and angular velocities are the time derivatives (central
    differences) of its own world poses when q moves with qd and link 0 moves
    rigidly with the given centre-of-mass and angular velocity."""
    from scipy.spatial.transform import Rotation

    from parity import link_coms, link_state_reference

    rob = S.robot(tree)
    o = _model(rob)
    rng = np.random.default_rng(40)
    B = 4
    q, qd, _, _ = S.samples(rob, B, 41)
    R0 = Rotation.random(B, random_state=42).as_matrix()
    p0, vc0, w0 = rng.uniform(-1, 1, (B, 3)), rng.normal(size=(B, 3)), rng.normal(size=(B, 3))
    c = link_coms(rob.urdf, rob.link_names)

    def root_at(t):
        Rt = Rotation.from_rotvec(t * w0).as_matrix() @ R0
        com = p0 + np.einsum("bij,j->bi", R0, c[0]) + t * vc0
        pt = com - np.einsum("bij,j->bi", Rt, c[0])
        xyzw = Rotation.from_matrix(Rt).as_quat()
        return np.concatenate([pt, xyzw[:, 3:], xyzw[:, :3], vc0, w0], 1)

    pw, Rw, vc, ww = link_state_reference(o, rob.urdf, rob.link_names, q, qd, root_at(0.0))
    h = 1e-6
    pp, Rp, _, _ = link_state_reference(o, rob.urdf, rob.link_names, q + h * qd, qd, root_at(h))
    pm, Rm, _, _ = link_state_reference(o, rob.urdf, rob.link_names, q - h * qd, qd, root_at(-h))
    com_p = pp + np.einsum("blij,lj->bli", Rp, c)
    com_m = pm + np.einsum("blij,lj->bli", Rm, c)
    assert np.allclose((com_p - com_m) / (2 * h), vc, atol=1e-6)
    dR = Rp @ np.swapaxes(Rm, -1, -2)  # exp(2 h [w])
    w_fd = Rotation.from_matrix(dR.reshape(-1, 3, 3)).as_rotvec().reshape(B, -1, 3) / (2 * h)
    assert np.allclose(w_fd, ww, atol=1e-6)
    # and the world poses are root * (model pose of link 0)^-1 * model pose
    assert np.allclose(pw[:, 0], root_at(0.0)[:, :3], atol=1e-12)
    assert np.allclose(Rw[:, 0], R0, atol=1e-12)
