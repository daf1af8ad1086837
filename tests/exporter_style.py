"""Rewrite a URDF into exactly what SAPIEN's exporter emits for the loaded
articulation (python/py_package/wrapper/urdf_exporter.py:21-152), as text:

* ``ET.tostring(..., encoding="utf8")`` output with its XML declaration, a
  ``<robot name="">`` element, a ``world`` link, then every link ``link_{i}``
  (i = link index, depth-first from the root), then per joint of child i the
  joint ``joint_{i}``, the dummy link ``link_dummy_{i}`` and the fixed joint
  ``joint_dummy_{i}`` (:138-142, :55-128);
* the root joint ``joint_{root}`` from ``world`` with no origin, ``fixed``
  (the exporter's ``force_fix_root``, :145-152);
* every joint (fixed ones included) from the parent link to the dummy link
  with origin = pose_in_parent, whose x axis is the joint axis (the SAPIEN URDF
  loader aligns x with the axis, python/py_package/wrapper/urdf_loader.py:510-573),
  ``<axis xyz="1 0 0"/>``, then a fixed joint to the child link with origin =
  pose_in_child^-1 (:74-126);
* revolute joints whose limits are beyond +-1e5 become ``continuous`` with
  limits clamped to +-1e5 (:64-67, :99-108);
* inertial = principal moments at the centre-of-mass pose whose rotation is
  the principal axes (``cmass_local_pose``, :21-47);
* FP32 values, printed as the exporter prints SAPIEN's float32 members:
  numpy float32 positions, angles, inertia and limits by ``str`` (shortest
  float32 repr) and the mass, a Python float holding a float32, by ``str``
  (``fp32=False`` prints the FP64 values instead: the same mechanism as the
  source URDF to rounding).

A link without an inertial element (SAPIEN computes one from its collision
shapes, not available here) is given mass 1 and inertia 1e-3 I at its origin.
``rewrite`` returns the text; ``export`` also returns the joint and link name
maps (source name -> exported name).
"""
import math
import xml.etree.ElementTree as ET

import numpy as np


def _rpy_of(R):
    """(roll, pitch, yaw) with R = Rz(yaw) Ry(pitch) Rx(roll); roll = 0 at gimbal lock."""
    s = -R[2, 0]
    if abs(s) >= 1.0 - 1e-12:
        p = math.copysign(math.pi / 2, s)
        return 0.0, p, math.atan2(-R[0, 1], R[1, 1])
    p = math.asin(s)
    return math.atan2(R[2, 1], R[2, 2]), p, math.atan2(R[1, 0], R[0, 0])


def _rpy_mat(r, p, y):
    cr, sr, cp, sp, cy, sy = math.cos(r), math.sin(r), math.cos(p), math.sin(p), math.cos(y), math.sin(y)
    return np.array([[cy, -sy, 0], [sy, cy, 0], [0, 0, 1]]) @ np.array([[cp, 0, sp], [0, 1, 0], [-sp, 0, cp]]) @ \
        np.array([[1, 0, 0], [0, cr, -sr], [0, sr, cr]])


def _rot(a, t):
    K = np.array([[0, -a[2], a[1]], [a[2], 0, -a[0]], [-a[1], a[0], 0]])
    return np.eye(3) + math.sin(t) * K + (1 - math.cos(t)) * K @ K


def _align_x_to(a, rng):
    """A rotation with R e_x = a (a random roll about a)."""
    a = a / np.linalg.norm(a)
    ex = np.array([1.0, 0, 0])
    v = np.cross(ex, a)
    s, c = np.linalg.norm(v), float(ex @ a)
    if s < 1e-12:
        R0 = np.eye(3) if c > 0 else _rot(np.array([0, 0, 1.0]), math.pi)
    else:
        R0 = _rot(v / s, math.atan2(s, c))
    return _rot(a, rng.uniform(-math.pi, math.pi)) @ R0


_FULL = [False]  # print full FP64 values instead (export(..., fp32=False))


def _f32(x):
    if _FULL[0]:
        return repr(float(x))
    return str(np.float32(x))  # numpy float32 member printed by str / an f-string


def _mass(x):
    return repr(float(x)) if _FULL[0] else str(float(np.float32(x)))  # Python float holding a float32


def _vec(v):
    return " ".join(_f32(x) for x in v)


def _floats(s, n, default):
    return [float(x) for x in s.split()] if s is not None else list(default)


def _origin(elem):
    o = elem.find("origin") if elem is not None else None
    if o is None:
        return np.zeros(3), np.eye(3)
    xyz = _floats(o.get("xyz"), 3, (0, 0, 0))
    rpy = _floats(o.get("rpy"), 3, (0, 0, 0))
    return np.array(xyz), _rpy_mat(*rpy)


def _principal(inertial):
    """(mass, com position, principal axes R, principal moments) of a URDF inertial."""
    if inertial is None:
        return 1.0, np.zeros(3), np.eye(3), np.full(3, 1e-3)
    c, Ro = _origin(inertial)
    m = float(inertial.find("mass").get("value"))
    ii = inertial.find("inertia")
    g = {k: float(ii.get(k, "0")) for k in ("ixx", "ixy", "ixz", "iyy", "iyz", "izz")}
    I = np.array([[g["ixx"], g["ixy"], g["ixz"]], [g["ixy"], g["iyy"], g["iyz"]], [g["ixz"], g["iyz"], g["izz"]]])
    I = Ro @ I @ Ro.T
    if np.allclose(I, np.diag(np.diag(I)), rtol=0, atol=1e-15):
        return m, c, np.eye(3), np.diag(I).copy()  # already principal: identity axes
    w, V = np.linalg.eigh(I)
    if np.linalg.det(V) < 0:
        V[:, 0] = -V[:, 0]
    return m, c, V, w


def export(urdf, seed=0, fp32=True):
    _FULL[0] = not fp32
    try:
        return _export(urdf, seed)
    finally:
        _FULL[0] = False


def _export(urdf, seed):
    rng = np.random.default_rng(seed)
    src = ET.fromstring(urdf)
    links = {l.get("name"): l for l in src.findall("link")}
    joints = src.findall("joint")
    by_child = {j.find("child").get("link"): j for j in joints}
    kids = {}
    for j in joints:
        kids.setdefault(j.find("parent").get("link"), []).append(j)
    root = [n for n in links if n not in by_child][0]
    order = []

    def visit(n):
        order.append(n)
        for j in kids.get(n, []):
            visit(j.find("child").get("link"))

    visit(root)
    index = {n: i for i, n in enumerate(order)}
    link_map = {n: f"link_{index[n]}" for n in order}
    joint_map = {}

    robot = ET.Element("robot", {"name": ""})
    ET.SubElement(robot, "link", {"name": "world"})
    for n in order:
        m, c, V, w = _principal(links[n].find("inertial"))
        el = ET.SubElement(robot, "link", {"name": link_map[n]})
        inl = ET.SubElement(el, "inertial")
        ET.SubElement(inl, "origin", {"xyz": _vec(c), "rpy": _vec(_rpy_of(V))})
        ET.SubElement(inl, "mass", {"value": _mass(m)})
        ET.SubElement(inl, "inertia", {"ixx": _f32(w[0]), "iyy": _f32(w[1]), "izz": _f32(w[2]),
                                       "ixy": "0", "ixz": "0", "iyz": "0"})
    rj = ET.SubElement(robot, "joint", {"name": f"joint_{index[root]}", "type": "fixed"})
    ET.SubElement(rj, "parent", {"link": "world"})
    ET.SubElement(rj, "child", {"link": link_map[root]})
    for n in order[1:]:
        j = by_child[n]
        i = index[n]
        jt = j.get("type")
        joint_map[j.get("name")] = f"joint_{i}"
        xyz, R = _origin(j)
        ax = j.find("axis")
        a = np.array(_floats(ax.get("xyz") if ax is not None else None, 3, (1, 0, 0)))
        Ra = _align_x_to(a, rng) if jt != "fixed" else np.eye(3)
        lim = j.find("limit")
        lo = float(lim.get("lower", "0")) if lim is not None else 0.0
        hi = float(lim.get("upper", "0")) if lim is not None else 0.0
        if jt == "continuous":
            lo, hi = -math.inf, math.inf  # SAPIEN loads it as a revolute joint without limits
            jt = "revolute"
        if jt == "revolute" and lo < -1e5 and hi > 1e5:
            jt = "continuous"
        ej = ET.SubElement(robot, "joint", {"name": f"joint_{i}", "type": jt})
        ET.SubElement(ej, "origin", {"xyz": _vec(xyz), "rpy": _vec(_rpy_of(R @ Ra))})
        ET.SubElement(ej, "axis", {"xyz": "1 0 0"})
        ET.SubElement(ej, "parent", {"link": link_map[j.find("parent").get("link")]})
        ET.SubElement(ej, "child", {"link": f"link_dummy_{i}"})
        if jt in ("prismatic", "revolute", "continuous"):
            lo32, hi32 = (np.float64(lo), np.float64(hi)) if _FULL[0] else (np.float32(lo), np.float32(hi))
            ET.SubElement(ej, "limit", {"effort": "0", "velocity": "0", "lower": str(max(lo32, -1e5)),
                                        "upper": str(min(hi32, 1e5))})
        ET.SubElement(robot, "link", {"name": f"link_dummy_{i}"})
        dj = ET.SubElement(robot, "joint", {"name": f"joint_dummy_{i}", "type": "fixed"})
        ET.SubElement(dj, "origin", {"xyz": _vec(np.zeros(3)), "rpy": _vec(_rpy_of(Ra.T))})
        ET.SubElement(dj, "axis", {"xyz": "0 0 0"})
        ET.SubElement(dj, "parent", {"link": f"link_dummy_{i}"})
        ET.SubElement(dj, "child", {"link": link_map[n]})
    return ET.tostring(robot, encoding="utf8").decode(), joint_map, link_map


def rewrite(urdf, seed=0, fp32=True):
    return export(urdf, seed, fp32)[0]
