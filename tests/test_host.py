"""CPU-only checks of the product's host side: the C ABI library loads and
exports every symbol include/sapien_amd.h declares, the C++ URDF reader and
model compiler agree with the oracle's independent Python builder, and the
drop-in class reproduces the reference's argument handling and error types
(pinocchio/pinocchio_model.cpp:33-36, 64-70, 99-140, 150-151)."""
import ctypes
import os
import xml.etree.ElementTree as ET

import numpy as np
import pytest

import sapien_amd as sa
from exporter_style import rewrite
from oracle.urdf_model import OracleModel
from sapien_amd import native
from sapien_amd import synthetic as S

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(native.LIB_PATH)
    syms = native.declared_symbols()
    assert len(syms) >= 25
    for s in syms:
        assert hasattr(lib, s), s
    assert b"gfx950" in native.lib().sapamd_version()


def _info(m):
    return m.nv, m.nq, m.nlinks


@pytest.mark.parametrize("name", list(S.ROBOTS))
def test_model_sizes_match_oracle(name):
    rob = S.robot(name)
    m = sa.PinocchioModel(rob.urdf, [0, 0, -9.81])
    o = OracleModel(rob.urdf)
    assert m.nv == o.nj == rob.nv
    assert m.nq == rob.nv + int(rob.continuous.sum())  # continuous joints are (cos, sin) in Pinocchio
    m.set_joint_order(rob.joint_names)
    m.set_link_order(rob.link_names)
    lo, hi = m.joint_limits()
    assert np.allclose(lo, rob.lower) and np.allclose(hi, rob.upper)


def test_fixtures_parse():
    for fname, nv, nq in (("movo_simple.urdf", 13, 20), ("panda.urdf", 9, 9)):
        text = open(os.path.join(GOLD, fname)).read()
        m = sa.PinocchioModel(text, [0, 0, -9.81])
        assert (m.nv, m.nq) == (nv, nq)
    m = sa.PinocchioModel(rewrite(S.torso30().urdf), [0, 0, -9.81])
    assert m.nv == 30


def test_reference_error_semantics():
    rob = S.chain7()
    m = sa.PinocchioModel(rob.urdf, [0, 0, -9.81])
    with pytest.raises(RuntimeError, match="posS2P failed"):  # compute before set_joint_order
        m.compute_forward_kinematics(np.zeros(7))
    with pytest.raises(ValueError, match="invalid names in setJointOrder"):
        m.set_joint_order(["joint_1", "nope"])
    with pytest.raises(RuntimeError, match="setJointOrder failed"):
        m.set_joint_order(rob.joint_names[:5])
    with pytest.raises(ValueError, match="invalid names in setLinkOrder"):
        m.set_link_order(["link_0", "nope"])
    m.set_joint_order(rob.joint_names)
    m.set_link_order(rob.link_names)
    with pytest.raises(RuntimeError, match="link index out of bound"):
        m.get_link_pose(8)
    with pytest.raises(RuntimeError, match="link index out of bound"):
        m.get_link_jacobian(100)
    with pytest.raises(RuntimeError, match="posS2P failed"):
        m.compute_inverse_dynamics(np.zeros(8), np.zeros(8), np.zeros(8))
    # before any compute, poses are the link placements (pinocchio::Data starts at identity)
    p = m.get_link_pose(3)
    assert np.allclose(p.p, 0) and np.allclose(p.q, [1, 0, 0, 0])
    assert np.all(m.get_link_jacobian(3) == 0)


def test_unsupported_joint_types_fail_at_compute():
    urdf = ('<robot name="r"><link name="world"/><link name="a"/><link name="b"/>'
            '<joint name="free" type="floating"><parent link="world"/><child link="a"/></joint>'
            '<joint name="j" type="revolute"><parent link="a"/><child link="b"/><limit lower="-1" upper="1"/></joint>'
            '</robot>')
    m = sa.PinocchioModel(urdf, [0, 0, -9.81])
    assert m.nv == 7 and m.nq == 8
    m.set_joint_order(["free", "j"])
    m.set_link_order(["a", "b"])
    with pytest.raises(RuntimeError, match="Unsupported joint in computation"):
        m.compute_forward_kinematics(np.zeros(7))


def test_bad_urdf_raises_value_error():
    with pytest.raises(ValueError):
        sa.PinocchioModel("<robot><link name='a'>", [0, 0, -9.81])
    with pytest.raises(ValueError):
        sa.PinocchioModel("<robot name='r'><link name='a'/><link name='b'/></robot>", [0, 0, -9.81])


def test_pose_type_matches_sapien_pose_kats():
    # unittest/test_sapien/test_math.py:27-42, 58-69, 79-95
    R = np.array([[-0.1301451, -0.5405447, 0.8311881, 0.1], [-0.9910400, 0.0963137, -0.0925388, 0.2],
                  [-0.0300335, -0.8357841, -0.5482362, 0.3], [0, 0, 0, 1]], dtype=np.float32)
    pose = sa.Pose(R)
    ref = np.array([-0.3232385, 0.5748429, -0.6660884, 0.3484233])
    assert np.allclose(pose.p, [0.1, 0.2, 0.3])
    assert np.allclose(pose.q, ref) or np.allclose(pose.q, -ref)
    p2 = sa.Pose()
    p2.rpy = [0.3, 0.4, 0.5]
    r = np.array([0.94628083, 0.09330659, 0.22656631, 0.21098383])
    assert np.allclose(p2.q, r) or np.allclose(p2.q, -r)
    assert np.allclose(p2.rpy, [0.3, 0.4, 0.5], atol=1e-6)
    T = sa.Pose([0.1, 0.2, 0.3], ref).to_transformation_matrix()
    assert np.allclose(T[:3, :3], R[:3, :3], atol=1e-6)
    a = sa.Pose([0.1, -0.2, 0.3], ref)
    b = a * a.inv()
    assert np.allclose(b.p, 0, atol=1e-6) and np.allclose(abs(b.q[0]), 1, atol=1e-6)
    import pickle
    c = pickle.loads(pickle.dumps(a))
    assert np.allclose(c.p, a.p) and np.allclose(c.q, a.q)


def test_condition_estimate_matches_oracle_mass_matrices():
    """The host cond(M) estimate that picks FP32 ABA or FP64 factorization:
    within the range the oracle's CRBA gives over random configurations, and
    far above the 1e4 switch for a chain with a 1e-5 kg tip."""
    import re

    import sapien_amd as sa
    from oracle.urdf_model import OracleModel

    for name in ("chain7", "quad18"):
        rob = S.robot(name)
        m = sa.PinocchioModel(rob.urdf, [0.0, 0.0, -9.81])
        m.set_joint_order(rob.joint_names)
        prec, cond = sa.dynamics_precision(m)
        o = OracleModel(rob.urdf)
        o.set_joint_order(rob.joint_names)
        ev = np.linalg.eigvalsh(o.crba(S.samples(rob, 512, 3)[0]))
        c = ev[:, -1] / ev[:, 0]
        assert np.median(c) <= cond <= 10 * c.max(), (name, cond, c.max())
        assert prec == ("fp64" if name == "quad18" else "fp32")  # switch above cond 1e4
    rob = S.chain7()
    i = rob.urdf.rfind("<inertial>")
    tail = re.sub(r'<mass value="[^"]*"/>', '<mass value="1e-05"/>', rob.urdf[i:], count=1)
    tail = re.sub(r'(ixx|iyy|izz)="[^"]*"', r'\1="1e-10"', tail, count=3)
    m = sa.PinocchioModel(rob.urdf[:i] + tail, [0.0, 0.0, -9.81])
    m.set_joint_order(rob.joint_names)
    prec, cond = sa.dynamics_precision(m)
    assert prec == "fp64" and cond > 1e6
    sa.set_dynamics_precision(m, "fp32")
    assert sa.dynamics_precision(m)[0] == "fp32"
    with pytest.raises(ValueError):
        sa.set_dynamics_precision(m, "fp16")


def test_philox_known_answers():
    """The numpy Philox4x32-10 the random-configuration tests check against,
    pinned by the generator's published known-answer vectors (Random123)."""
    from philox import philox4x32_10

    kats = [((0, 0, 0, 0), (0, 0), (0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8)),
            ((0xffffffff,) * 4, (0xffffffff, 0xffffffff), (0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd)),
            ((0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344), (0xa4093822, 0x299f31d0),
             (0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1))]
    for c, k, ref in kats:
        assert tuple(int(x) for x in philox4x32_10(*c, *k)) == ref
