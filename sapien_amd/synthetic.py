"""Seeded synthetic robots and input samples for the articulated-body workloads.

Four trees, written as URDF text in the layout SAPIEN's exporter produces
(``link_{i}`` / ``joint_{i}`` named by child-link index,
``python/py_package/wrapper/urdf_exporter.py:21-152``) so they load through the
drop-in ``PinocchioModel`` exactly like an exported articulation:

========  ====  ======================================================================
tree      nv    structure
========  ====  ======================================================================
chain7    7     serial revolute arm, 8 links (fixed base ``link_0``)
arm9      9     chain7 + two prismatic fingers (``+y``/``-y``, [0, 0.04] m) off link 7
quad18    18    floating base emulated as P-P-P-R-R-R through 5 massless virtual links,
                trunk, 4 legs x 3 revolute, hips at (+-0.2, +-0.2, 0) m; 13 user links
torso30   30    6 continuous wheels + 3 revolute torso + two arm9-style arms + 3-dof head;
                31 links
========  ====  ======================================================================

Per link (seeded ``numpy.random.default_rng(seed)``): outgoing offset length
U[0.05, 0.4] m along a uniform random unit vector (normalized N(0, I3)); joint
axis uniform over {+-x, +-y, +-z}; mass U[0.5, 3] kg; diagonal principal
inertia U[1e-3, 1e-1] kg m^2 (link-aligned); centre of mass at the midpoint of
the outgoing offset; revolute limits +-2.8 rad. Default seeds: chain7 0, arm9 1,
quad18 2, torso30 3.

Samples (seed 1000 + workload index): q ~ U(limits) (continuous joints
U(-pi, pi)), qd ~ N(0, 1), qdd ~ N(0, var 4), tau ~ N(0, var 10), rollout
torques ~ N(0, var 5) -- N(mu, var) with the variance as the second argument.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

AXES = np.array([[1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1]], dtype=float)


@dataclass
class Robot:
    name: str
    urdf: str
    joint_names: list
    link_names: list
    lower: np.ndarray
    upper: np.ndarray
    continuous: np.ndarray  # bool per user dof
    ee_link: int  # user index of the default end-effector link (deepest link of the first branch)
    meta: dict = field(default_factory=dict)

    @property
    def nv(self):
        return len(self.joint_names)

    @property
    def nlinks(self):
        return len(self.link_names)


def _fmt(v):
    return " ".join(repr(float(x)) for x in v)


class _Builder:
    def __init__(self, seed):
        self.rng = np.random.default_rng(seed)
        self.links = []  # (name, mass, com, diag)
        self.joints = []  # (name, type, parent, child, xyz, axis, lo, hi)

    def offset(self):
        ell = self.rng.uniform(0.05, 0.4)
        u = self.rng.normal(size=3)
        return ell * u / np.linalg.norm(u)

    def axis(self):
        return AXES[self.rng.integers(0, 6)]

    def link(self, name, out_offset=None, massless=False):
        if massless:
            self.links.append((name, 0.0, np.zeros(3), np.zeros(3)))
            return
        m = self.rng.uniform(0.5, 3.0)
        d = self.rng.uniform(1e-3, 1e-1, size=3)
        com = 0.5 * out_offset if out_offset is not None else np.zeros(3)
        self.links.append((name, m, com, d))

    def joint(self, name, jtype, parent, child, xyz, axis, lo=-2.8, hi=2.8):
        self.joints.append((name, jtype, parent, child, np.asarray(xyz, float), np.asarray(axis, float), lo, hi))

    def urdf(self, robot_name):
        out = [f'<robot name="{robot_name}">']
        for name, m, com, d in self.links:
            if m == 0.0:
                out.append(f'  <link name="{name}"/>')
                continue
            out.append(f'  <link name="{name}">\n    <inertial>\n      <origin xyz="{_fmt(com)}" rpy="0 0 0"/>'
                       f'\n      <mass value="{float(m)!r}"/>\n      <inertia ixx="{float(d[0])!r}" ixy="0" ixz="0" iyy="{float(d[1])!r}"'
                       f' iyz="0" izz="{float(d[2])!r}"/>\n    </inertial>\n  </link>')
        for name, jt, parent, child, xyz, axis, lo, hi in self.joints:
            s = (f'  <joint name="{name}" type="{jt}">\n    <origin xyz="{_fmt(xyz)}" rpy="0 0 0"/>\n'
                 f'    <axis xyz="{_fmt(axis)}"/>\n    <parent link="{parent}"/>\n    <child link="{child}"/>\n')
            if jt in ("revolute", "prismatic"):
                s += f'    <limit effort="0" velocity="0" lower="{float(lo)!r}" upper="{float(hi)!r}"/>\n'
            s += "  </joint>"
            out.append(s)
        out.append("</robot>")
        return "\n".join(out)


def _arm(b, base_link, first_index, n_rev, fingers, base_offset):
    """Serial revolute chain of n_rev links hanging off base_link at base_offset,
    optionally ending in two prismatic fingers. Returns (indices, tip_offset)."""
    offsets = [base_offset]
    idx = []
    parent = base_link
    off_in = base_offset
    for k in range(n_rev):
        i = first_index + k
        out_off = b.offset()
        b.link(f"link_{i}", out_off)
        b.joint(f"joint_{i}", "revolute", parent, f"link_{i}", off_in, b.axis())
        parent, off_in = f"link_{i}", out_off
        idx.append(i)
        offsets.append(out_off)
    if fingers:
        for s, sign in enumerate((1.0, -1.0)):
            i = first_index + n_rev + s
            b.link(f"link_{i}", b.offset())
            b.joint(f"joint_{i}", "prismatic", parent, f"link_{i}", off_in, [0.0, sign, 0.0], 0.0, 0.04)
            idx.append(i)
    return idx


def _finish(b, name, link_idx, joint_idx, ee_user, meta=None):
    urdf = b.urdf(name)
    jmap = {j[0]: j for j in b.joints}
    jn = [f"joint_{i}" for i in joint_idx]
    lo = np.array([jmap[n][6] if jmap[n][1] != "continuous" else -math.pi for n in jn])
    hi = np.array([jmap[n][7] if jmap[n][1] != "continuous" else math.pi for n in jn])
    cont = np.array([jmap[n][1] == "continuous" for n in jn])
    return Robot(name, urdf, jn, [f"link_{i}" for i in link_idx], lo, hi, cont, ee_user, meta or {})


def chain7(seed=0):
    b = _Builder(seed)
    b.link("link_0", None)
    first = b.offset()
    idx = _arm(b, "link_0", 1, 7, False, first)
    return _finish(b, "chain7", [0] + idx, idx, ee_user=7)


def chain(n, seed=0):
    """Serial revolute chain of n joints (n + 1 links), generated like chain7."""
    b = _Builder(seed)
    b.link("link_0", None)
    first = b.offset()
    idx = _arm(b, "link_0", 1, n, False, first)
    return _finish(b, f"chain{n}", [0] + idx, idx, ee_user=n)


def arm9(seed=1):
    b = _Builder(seed)
    b.link("link_0", None)
    first = b.offset()
    idx = _arm(b, "link_0", 1, 7, True, first)
    return _finish(b, "arm9", [0] + idx, idx, ee_user=7)


def quad18(seed=2):
    b = _Builder(seed)
    b.link("link_0", massless=True)
    for i in range(1, 6):
        b.link(f"link_{i}", massless=True)
    trunk_out = b.offset()
    b.link("link_6", trunk_out)
    z = [0.0, 0.0, 0.0]
    for i, ax in enumerate(([1, 0, 0], [0, 1, 0], [0, 0, 1])):
        b.joint(f"joint_{i + 1}", "prismatic", f"link_{i}", f"link_{i + 1}", z, ax, -1.0, 1.0)
    for i, ax in enumerate(([1, 0, 0], [0, 1, 0], [0, 0, 1])):
        b.joint(f"joint_{i + 4}", "revolute", f"link_{i + 3}", f"link_{i + 4}", z, ax)
    legs = []
    nxt = 7
    for hx, hy in ((0.2, 0.2), (0.2, -0.2), (-0.2, 0.2), (-0.2, -0.2)):
        legs += _arm(b, "link_6", nxt, 3, False, np.array([hx, hy, 0.0]))
        nxt += 3
    joints = list(range(1, 7)) + legs
    return _finish(b, "quad18", [6] + legs, joints, ee_user=3, meta={"virtual_links": 5})


def torso30(seed=3):
    b = _Builder(seed)
    b.link("link_0", b.offset())
    for i in range(1, 7):  # wheels
        out = b.offset()
        b.link(f"link_{i}", out)
        b.joint(f"joint_{i}", "continuous", "link_0", f"link_{i}", b.offset(), b.axis())
    torso = _arm(b, "link_0", 7, 3, False, b.offset())
    tip = b.offset()
    left = _arm(b, "link_9", 10, 7, True, tip)
    right = _arm(b, "link_9", 19, 7, True, b.offset())
    head = _arm(b, "link_9", 28, 3, False, b.offset())
    idx = list(range(1, 31))
    return _finish(b, "torso30", [0] + idx, idx, ee_user=16)


ROBOTS = {"chain7": chain7, "arm9": arm9, "quad18": quad18, "torso30": torso30}


def robot(name, seed=None):
    f = ROBOTS[name]
    return f() if seed is None else f(seed)


def samples(rob, B, seed, dtype=np.float64):
    """Seeded (q, qd, qdd, tau) of shape (B, nv): q ~ U(limits), qd ~ N(0,1),
    qdd ~ N(0, var 4), tau ~ N(0, var 10)."""
    rng = np.random.default_rng(seed)
    n = rob.nv
    q = rng.uniform(rob.lower, rob.upper, size=(B, n))
    qd = rng.normal(0.0, 1.0, size=(B, n))
    qdd = rng.normal(0.0, 2.0, size=(B, n))
    tau = rng.normal(0.0, math.sqrt(10.0), size=(B, n))
    return [x.astype(dtype) for x in (q, qd, qdd, tau)]


# ---------------------------------------------------------------- workloads
# The mixed batch: chain7, arm9, quad18 and torso30 in the ratio 4:3:2:1.
MIX = (("chain7", 4), ("arm9", 3), ("quad18", 2), ("torso30", 1))


def split_mix(total):
    """total split 4:3:2:1 over the four trees (rounded, summing to total)."""
    w = [k for _, k in MIX]
    n = [round(total * k / sum(w)) for k in w]
    n[0] += total - sum(n)
    return n


ROLLOUT_STEPS, ROLLOUT_CKPT, ROLLOUT_DT = 128, 16, 1.0 / 240.0


def rollout_workload(rob, B, seed, T=ROLLOUT_STEPS, dtype=np.float32):
    """Inputs of B rollouts: q0, qd0 from samples(rob, B, seed) and torques
    tau (B, T, nv) ~ N(0, var 5) from default_rng(seed + 100)."""
    q0, qd0, _, _ = samples(rob, B, seed, dtype=dtype)
    tau = np.random.default_rng(seed + 100).normal(0.0, math.sqrt(5.0), size=(B, T, rob.nv)).astype(dtype)
    return q0, qd0, tau


def gradient_workload(rob, B, dtype=np.float32):
    """Inputs of the pose + torque loss over B environments: (q, qd, qdd, tau*)
    from samples(rob, B, 1003) and the configurations q* (seed 1013) whose
    end-effector positions FK(q*) are the targets."""
    q, qd, qdd, tau_star = samples(rob, B, 1003, dtype=dtype)
    q_star = samples(rob, B, 1013, dtype=dtype)[0]
    return q, qd, qdd, tau_star, q_star
