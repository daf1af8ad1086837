"""ORACLE — test infrastructure only.

Only ``tests/``, ``__graft_entry__.smoke()`` and the CPU-baseline leg of
``bench.py`` may import this module. The product (``sapien_amd``) never does:
it has its own C++ URDF parser and model compiler, and these two independent
implementations are checked against each other.

This module turns URDF text into the FP64 model tables consumed by the C++
restatement in ``rbd_oracle.hpp`` and wraps that library through ctypes with
the reference's method semantics.

URDF -> model follows Pinocchio v2.6.18's URDF parser as called at
``pinocchio/pinocchio_model.cpp:42`` (``pinocchio::urdf::buildModelFromXML``),
restated from its documented behaviour:

* the root link (the one link that is no joint's child) becomes a BODY frame on
  the universe joint; its inertia is attached to the universe and never moves;
* children are visited depth-first, in urdfdom's order, which is the order of a
  ``std::map`` keyed by joint name (byte-wise sorted names);
* a movable joint gets placement ``parent_frame.placement * origin`` relative to
  the parent frame's joint, and its child link's BODY frame sits on the new joint
  with identity placement; the child's inertia is attached to the new joint;
* a fixed joint creates no joint: the child link's BODY frame sits on the parent
  frame's joint at ``parent_frame.placement * origin`` and its inertia is merged
  into that joint (``Inertia::operator+=``, parallel-axis theorem);
* ``<inertial><origin rpy>`` rotates the inertia tensor: ``I = R I_urdf R^T``;
* joint axes are normalized; ``revolute``/``continuous``/``prismatic`` are the
  1-dof joints supported by ``posS2P`` (``pinocchio_model.cpp:49-72``);
  ``continuous`` is stored as nq = 2 (cos, sin) by Pinocchio and as an angle here.

Users index dofs and links in the order given to ``set_joint_order`` /
``set_link_order`` (``pinocchio_model.cpp:99-140``).
"""
from __future__ import annotations

import ctypes
import math
import os
import xml.etree.ElementTree as ET
from dataclasses import dataclass, field

import numpy as np

REVOLUTE, PRISMATIC, CONTINUOUS = 0, 1, 2
_JTYPES = {"revolute": REVOLUTE, "prismatic": PRISMATIC, "continuous": CONTINUOUS}

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


def _lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle_rbd.so")
        if not os.path.exists(path):
            raise RuntimeError(f"oracle library not built: run `make -C {_HERE}`")
        lib = ctypes.CDLL(path)
        P = ctypes.c_void_p
        I64 = ctypes.c_int64
        D = ctypes.c_double
        lib.oracle_create.restype = P
        lib.oracle_destroy.argtypes = [P]
        lib.oracle_clik.restype = ctypes.c_int
        lib.oracle_fk_jac_rnea_timed.restype = D
        for name in ["oracle_fk_poses", "oracle_jacobian_world", "oracle_crba"]:
            getattr(lib, name).restype = None
        _LIB = lib
    return _LIB


def _ptr(a, ctype=ctypes.c_double):
    return a.ctypes.data_as(ctypes.POINTER(ctype))


# --------------------------------------------------------------------------- geometry
def rpy_to_matrix(r, p, y):
    """URDF fixed-axis roll-pitch-yaw: R = Rz(y) Ry(p) Rx(r)."""
    cr, sr, cp, sp, cy, sy = math.cos(r), math.sin(r), math.cos(p), math.sin(p), math.cos(y), math.sin(y)
    Rx = np.array([[1, 0, 0], [0, cr, -sr], [0, sr, cr]])
    Ry = np.array([[cp, 0, sp], [0, 1, 0], [-sp, 0, cp]])
    Rz = np.array([[cy, -sy, 0], [sy, cy, 0], [0, 0, 1]])
    return Rz @ Ry @ Rx


def _floats(s, n, default):
    if s is None:
        return list(default)
    v = [float(x) for x in s.split()]
    if len(v) != n:
        raise ValueError(f"expected {n} numbers, got {s!r}")
    return v


def _origin(elem):
    T = np.eye(4)
    if elem is None:
        return T
    xyz = _floats(elem.get("xyz"), 3, (0, 0, 0))
    rpy = _floats(elem.get("rpy"), 3, (0, 0, 0))
    T[:3, :3] = rpy_to_matrix(*rpy)
    T[:3, 3] = xyz
    return T


@dataclass
class Inertia:
    """Additive form: mass, first moment h = m c, rotational inertia about the frame origin."""

    m: float = 0.0
    h: np.ndarray = field(default_factory=lambda: np.zeros(3))
    IO: np.ndarray = field(default_factory=lambda: np.zeros((3, 3)))

    @staticmethod
    def from_com(m, c, Ic):
        c = np.asarray(c, float)
        C = _skew(c)
        return Inertia(m, m * c, np.asarray(Ic, float) - m * C @ C)

    def transformed(self, T):
        """Express in the parent frame of placement T (child -> parent)."""
        R, p = T[:3, :3], T[:3, 3]
        m, c, Ic = self.com_form()
        return Inertia.from_com(m, R @ c + p, R @ Ic @ R.T)

    def __add__(self, o):
        return Inertia(self.m + o.m, self.h + o.h, self.IO + o.IO)

    def com_form(self):
        c = self.h / max(self.m, np.finfo(float).eps) if self.m > 0 else np.zeros(3)
        C = _skew(c)
        return self.m, c, self.IO + self.m * C @ C


def _skew(v):
    return np.array([[0, -v[2], v[1]], [v[2], 0, -v[0]], [-v[1], v[0], 0]], dtype=float)


# --------------------------------------------------------------------------- URDF
@dataclass
class UJoint:
    name: str
    type: str
    parent: str
    child: str
    origin: np.ndarray
    axis: np.ndarray
    lower: float = 0.0
    upper: float = 0.0


def parse_urdf(text):
    root = ET.fromstring(text)
    if root.tag != "robot":
        raise ValueError("URDF root element must be <robot>")
    links, joints = {}, {}
    for e in root.findall("link"):
        inert = Inertia()
        ie = e.find("inertial")
        if ie is not None:
            T = _origin(ie.find("origin"))
            me = ie.find("mass")
            m = float(me.get("value")) if me is not None else 0.0
            ii = ie.find("inertia")
            I = np.zeros((3, 3))
            if ii is not None:
                g = lambda k: float(ii.get(k, "0"))
                I = np.array([[g("ixx"), g("ixy"), g("ixz")], [g("ixy"), g("iyy"), g("iyz")], [g("ixz"), g("iyz"), g("izz")]])
            R = T[:3, :3]
            inert = Inertia.from_com(m, T[:3, 3], R @ I @ R.T)
        links[e.get("name")] = inert
    for e in root.findall("joint"):
        jt = e.get("type")
        ax = e.find("axis")
        axis = np.array(_floats(ax.get("xyz") if ax is not None else None, 3, (1, 0, 0)))
        lim = e.find("limit")
        lo = float(lim.get("lower", "0")) if lim is not None else 0.0
        hi = float(lim.get("upper", "0")) if lim is not None else 0.0
        joints[e.get("name")] = UJoint(e.get("name"), jt, e.find("parent").get("link"), e.find("child").get("link"),
                                       _origin(e.find("origin")), axis, lo, hi)
    return links, joints


class OracleModel:
    """FP64 oracle with the reference PinocchioModel's semantics (user order I/O)."""

    def __init__(self, urdf_text, gravity=(0.0, 0.0, -9.81)):
        links, joints = parse_urdf(urdf_text)
        self.gravity = np.asarray(gravity, float)
        children = {ln: [] for ln in links}
        child_set = set()
        for jn in sorted(joints):  # std::map<std::string, Joint> order
            j = joints[jn]
            children[j.parent].append(j)
            child_set.add(j.child)
        roots = [ln for ln in links if ln not in child_set]
        if len(roots) != 1:
            raise ValueError(f"URDF must have exactly one root link, got {roots}")
        self.root_link = roots[0]
        # frames: link name -> (joint index or -1, placement 4x4)
        self.frames = {self.root_link: (-1, np.eye(4))}
        self.jnames, self.jparent, self.jtype, self.jaxis, self.jplace, self.jinertia = [], [], [], [], [], []
        self.jlimits = []
        self.unsupported = []

        def visit(link):
            for j in children[link]:
                pj, pplace = self.frames[link]
                place = pplace @ j.origin
                if j.type == "fixed":
                    self.frames[j.child] = (pj, place)
                    if pj >= 0:
                        self.jinertia[pj] = self.jinertia[pj] + links[j.child].transformed(place)
                else:
                    if j.type not in _JTYPES:
                        self.unsupported.append(j.name)
                        raise ValueError(f"joint type {j.type} is not supported by the oracle")
                    idx = len(self.jnames)
                    self.jnames.append(j.name)
                    self.jparent.append(pj)
                    self.jtype.append(_JTYPES[j.type])
                    n = np.linalg.norm(j.axis)
                    self.jaxis.append(j.axis / n)
                    self.jplace.append(place)
                    self.jinertia.append(links[j.child])
                    self.jlimits.append((j.lower, j.upper))
                    self.frames[j.child] = (idx, np.eye(4))
                visit(j.child)

        visit(self.root_link)
        self.nj = len(self.jnames)
        self.vuser = list(range(self.nj))
        self.link_names = []
        self._h = None

    # ------------------------------------------------------------------ orders
    def set_joint_order(self, names):
        v = []
        for nm in names:
            if nm not in self.jnames:
                raise ValueError(f"invalid joint name {nm}")
            v.append(self.jnames.index(nm))
        if len(v) != self.nj or len(set(v)) != self.nj:
            raise RuntimeError("setJointOrder failed")
        self.vuser = [0] * self.nj
        for u, j in enumerate(v):
            self.vuser[j] = u
        self._reset()

    def set_link_order(self, names):
        for nm in names:
            if nm not in self.frames:
                raise ValueError(f"invalid link name {nm}")
        self.link_names = list(names)
        self._reset()

    def _reset(self):
        if self._h is not None:
            _lib().oracle_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self._reset()
        except Exception:
            pass

    @property
    def nv(self):
        return self.nj

    @property
    def nlinks(self):
        return len(self.link_names)

    def tables(self):
        nj, L = self.nj, len(self.link_names)
        jplace = np.zeros((nj, 12))
        inertia = np.zeros((nj, 10))
        for j in range(nj):
            T = self.jplace[j]
            jplace[j, :9] = T[:3, :3].reshape(-1)
            jplace[j, 9:] = T[:3, 3]
            m, c, Ic = self.jinertia[j].com_form()
            inertia[j] = [m, *c, Ic[0, 0], Ic[0, 1], Ic[0, 2], Ic[1, 1], Ic[1, 2], Ic[2, 2]]
        link_joint = np.zeros(L, np.int32)
        link_place = np.zeros((L, 12))
        for l, nm in enumerate(self.link_names):
            j, T = self.frames[nm]
            link_joint[l] = j
            link_place[l, :9] = T[:3, :3].reshape(-1)
            link_place[l, 9:] = T[:3, 3]
        return dict(
            parent=np.asarray(self.jparent, np.int32),
            jtype=np.asarray(self.jtype, np.int32),
            axis=np.asarray(self.jaxis, float).reshape(nj, 3),
            jplace=jplace,
            inertia=inertia,
            vuser=np.asarray(self.vuser, np.int32),
            link_joint=link_joint,
            link_place=link_place,
        )

    def handle(self):
        if self._h is None:
            t = self.tables()
            self._keep = t
            self._grav = np.ascontiguousarray(self.gravity, float)
            I = ctypes.c_int
            self._h = _lib().oracle_create(
                I(self.nj), _ptr(t["parent"], I), _ptr(t["jtype"], I), _ptr(t["axis"]), _ptr(t["jplace"]),
                _ptr(t["inertia"]), _ptr(t["vuser"], I), I(len(self.link_names)), _ptr(t["link_joint"], I),
                _ptr(t["link_place"]), _ptr(self._grav))
        return ctypes.c_void_p(self._h)

    # ------------------------------------------------------------------ helpers
    def _q(self, q):
        q = np.ascontiguousarray(np.asarray(q, float))
        if q.ndim == 1:
            q = q[None]
        if q.shape[-1] != self.nj:
            raise RuntimeError("posS2P failed")
        return q

    def limits(self):
        lo = np.zeros(self.nj)
        hi = np.zeros(self.nj)
        for j in range(self.nj):
            u = self.vuser[j]
            if self.jtype[j] == CONTINUOUS:
                lo[u], hi[u] = -math.pi, math.pi
            else:
                lo[u], hi[u] = self.jlimits[j]
        return lo, hi

    # ------------------------------------------------------------------ algorithms (batched, user order)
    def link_poses(self, q):
        q = self._q(q)
        out = np.zeros((q.shape[0], self.nlinks, 7))
        _lib().oracle_fk_poses(self.handle(), ctypes.c_int64(q.shape[0]), _ptr(q), _ptr(out))
        return out

    def link_mats(self, q):
        q = self._q(q)
        B, L = q.shape[0], self.nlinks
        R = np.zeros((B, L, 3, 3))
        p = np.zeros((B, L, 3))
        _lib().oracle_fk_mats(self.handle(), ctypes.c_int64(B), _ptr(q), _ptr(R), _ptr(p))
        return R, p

    def jacobian_world(self, q):
        q = self._q(q)
        out = np.zeros((q.shape[0], 6, self.nj))
        _lib().oracle_jacobian_world(self.handle(), ctypes.c_int64(q.shape[0]), _ptr(q), _ptr(out))
        return out

    def link_jacobian(self, q, link, local=False):
        q = self._q(q)
        out = np.zeros((q.shape[0], 6, self.nj))
        _lib().oracle_link_jacobian(self.handle(), ctypes.c_int64(q.shape[0]), _ptr(q), ctypes.c_int(link),
                                    ctypes.c_int(int(local)), _ptr(out))
        return out

    def rnea(self, q, qd, qdd):
        q, qd, qdd = self._q(q), self._q(qd), self._q(qdd)
        out = np.zeros_like(q)
        _lib().oracle_rnea(self.handle(), ctypes.c_int64(q.shape[0]), _ptr(q), _ptr(qd), _ptr(qdd), _ptr(out))
        return out

    def crba(self, q):
        q = self._q(q)
        out = np.zeros((q.shape[0], self.nj, self.nj))
        _lib().oracle_crba(self.handle(), ctypes.c_int64(q.shape[0]), _ptr(q), _ptr(out))
        return out

    def aba(self, q, qd, tau):
        q, qd, tau = self._q(q), self._q(qd), self._q(tau)
        out = np.zeros_like(q)
        _lib().oracle_aba(self.handle(), ctypes.c_int64(q.shape[0]), _ptr(q), _ptr(qd), _ptr(tau), _ptr(out))
        return out

    def coriolis(self, q, qd):
        q, qd = self._q(q), self._q(qd)
        out = np.zeros((q.shape[0], self.nj, self.nj))
        _lib().oracle_coriolis(self.handle(), ctypes.c_int64(q.shape[0]), _ptr(q), _ptr(qd), _ptr(out))
        return out

    def rnea_derivs(self, q, qd, qdd):
        q, qd, qdd = self._q(q), self._q(qd), self._q(qdd)
        B, n = q.shape
        outs = [np.zeros((B, n, n)) for _ in range(3)]
        _lib().oracle_rnea_derivs(self.handle(), ctypes.c_int64(B), _ptr(q), _ptr(qd), _ptr(qdd),
                                  *[_ptr(o) for o in outs])
        return outs

    def link_pos_jac(self, q, link):
        q = self._q(q)
        out = np.zeros((q.shape[0], 3, self.nj))
        _lib().oracle_link_pos_jac(self.handle(), ctypes.c_int64(q.shape[0]), _ptr(q), ctypes.c_int(link), _ptr(out))
        return out

    def clik(self, link, target7, q0=None, mask=None, eps=1e-4, max_iterations=1000, dt=0.1, damp=1e-6):
        t = np.ascontiguousarray(np.asarray(target7, float))
        qo = np.zeros(self.nj)
        err = np.zeros(6)
        ok = ctypes.c_int(0)
        q0p = _ptr(np.ascontiguousarray(np.asarray(q0, float))) if q0 is not None else None
        if mask is not None:
            mk = np.ascontiguousarray(np.asarray(mask, np.int32))
            mp = _ptr(mk, ctypes.c_int)
        else:
            mp = None
        if q0 is not None:
            q0a = np.ascontiguousarray(np.asarray(q0, float))
            q0p = _ptr(q0a)
        it = _lib().oracle_clik(self.handle(), ctypes.c_int(link), _ptr(t), q0p, mp, ctypes.c_double(eps),
                                ctypes.c_int(max_iterations), ctypes.c_double(dt), ctypes.c_double(damp),
                                _ptr(qo), ctypes.byref(ok), _ptr(err))
        return qo, bool(ok.value), err, it

    def rollout(self, q0, qd0, tau, dt):
        q0, qd0 = self._q(q0), self._q(qd0)
        tau = np.ascontiguousarray(np.asarray(tau, float))
        B, T = tau.shape[0], tau.shape[1]
        qT, qdT = np.zeros_like(q0), np.zeros_like(q0)
        _lib().oracle_rollout(self.handle(), ctypes.c_int64(B), ctypes.c_int(T), ctypes.c_double(dt), _ptr(q0),
                              _ptr(qd0), _ptr(tau), _ptr(qT), _ptr(qdT))
        return qT, qdT

    def rollout_jvp(self, q0, qd0, tau, dq0, dqd0, dtau, dt):
        a = [np.ascontiguousarray(np.asarray(x, float)) for x in (q0, qd0, tau, dq0, dqd0, dtau)]
        T = a[2].shape[0]
        dqT, dqdT = np.zeros(self.nj), np.zeros(self.nj)
        _lib().oracle_rollout_jvp(self.handle(), ctypes.c_int(T), ctypes.c_double(dt), *[_ptr(x) for x in a],
                                  _ptr(dqT), _ptr(dqdT))
        return dqT, dqdT

    def fk_jac_rnea_timed(self, q, qd, qdd, threads=1, out=None):
        """Times the C++ loop only; pass out=(poses, J, tau) from an earlier call
        to reuse (already touched) output memory across repeated timings."""
        q, qd, qdd = self._q(q), self._q(qd), self._q(qdd)
        B = q.shape[0]
        if out is None:
            out = (np.zeros((B, self.nlinks, 7)), np.zeros((B, 6, self.nj)), np.zeros((B, self.nj)))
        poses, J, tau = out
        secs = _lib().oracle_fk_jac_rnea_timed(self.handle(), ctypes.c_int64(B), _ptr(q), _ptr(qd), _ptr(qdd),
                                               _ptr(poses), _ptr(J), _ptr(tau), ctypes.c_int(threads))
        return secs, poses, J, tau

    def opcount_fk_jac_rnea(self, q, qd, qdd):
        a = [np.ascontiguousarray(np.asarray(x, float)) for x in (q, qd, qdd)]
        out = np.zeros(5, np.uint64)
        _lib().oracle_opcount_fk_jac_rnea(self.handle(), *[_ptr(x) for x in a], _ptr(out, ctypes.c_uint64))
        return dict(zip(["adds", "muls", "divs", "sqrts", "trig"], out.tolist()))


def log6(pose7, with_jac=True):
    pose7 = np.ascontiguousarray(np.asarray(pose7, float))
    o6 = np.zeros(6)
    o36 = np.zeros(36)
    _lib().oracle_log6(_ptr(pose7), _ptr(o6), _ptr(o36) if with_jac else None)
    return o6, o36.reshape(6, 6)


def log6_mat(R, p):
    R = np.ascontiguousarray(np.asarray(R, float))
    p = np.ascontiguousarray(np.asarray(p, float))
    o6 = np.zeros(6)
    o36 = np.zeros(36)
    _lib().oracle_log6_mat(_ptr(R), _ptr(p), _ptr(o6), _ptr(o36))
    return o6, o36.reshape(6, 6)
