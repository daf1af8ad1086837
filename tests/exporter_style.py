"""Rewrite a URDF into the layout SAPIEN's exporter emits
(python/py_package/wrapper/urdf_exporter.py:52-152): a ``world`` root, every
movable joint from the parent link to ``link_dummy_i`` with its axis along +x of
a rotated joint frame (the SAPIEN URDF loader aligns x with the axis,
python/py_package/wrapper/urdf_loader.py:510-573), then a fixed ``joint_dummy_i``
back to ``link_i`` with origin = pose_in_child^-1. The rewritten robot is the
same mechanism, so every model output must be unchanged.
"""
import math
import xml.etree.ElementTree as ET

import numpy as np


def _rpy_of(R):
    p = math.asin(max(-1.0, min(1.0, -R[2, 0])))
    r = math.atan2(R[2, 1], R[2, 2])
    y = math.atan2(R[1, 0], R[0, 0])
    return r, p, y


def _rpy_mat(r, p, y):
    cr, sr, cp, sp, cy, sy = math.cos(r), math.sin(r), math.cos(p), math.sin(p), math.cos(y), math.sin(y)
    return np.array([[cy, -sy, 0], [sy, cy, 0], [0, 0, 1]]) @ np.array([[cp, 0, sp], [0, 1, 0], [-sp, 0, cp]]) @ \
        np.array([[1, 0, 0], [0, cr, -sr], [0, sr, cr]])


def _rot(a, t):
    K = np.array([[0, -a[2], a[1]], [a[2], 0, -a[0]], [-a[1], a[0], 0]])
    return np.eye(3) + math.sin(t) * K + (1 - math.cos(t)) * K @ K


def _align_x_to(a, rng):
    """A rotation with R e_x = a whose rpy round-trips (avoids the pitch = +-90 deg gimbal)."""
    a = a / np.linalg.norm(a)
    ex = np.array([1.0, 0, 0])
    v = np.cross(ex, a)
    s, c = np.linalg.norm(v), float(ex @ a)
    if s < 1e-12:
        R0 = np.eye(3) if c > 0 else _rot(np.array([0, 0, 1.0]), math.pi)
    else:
        R0 = _rot(v / s, math.atan2(s, c))
    for _ in range(100):
        R = _rot(a, rng.uniform(-math.pi, math.pi)) @ R0
        if np.allclose(_rpy_mat(*_rpy_of(R)), R, atol=1e-12) and np.allclose(_rpy_mat(*_rpy_of(R.T)), R.T, atol=1e-12):
            return R
    raise RuntimeError("no gimbal-free alignment found")


def _fmt(v):
    return " ".join(repr(float(x)) for x in v)


def rewrite(urdf, seed=0):
    rng = np.random.default_rng(seed)
    src = ET.fromstring(urdf)
    out = ET.Element("robot", {"name": ""})
    ET.SubElement(out, "link", {"name": "world"})
    child_links = {j.find("child").get("link") for j in src.findall("joint")}
    for l in src.findall("link"):
        out.append(l)
    root = [l.get("name") for l in src.findall("link") if l.get("name") not in child_links][0]
    rj = ET.SubElement(out, "joint", {"name": "joint_root", "type": "fixed"})
    ET.SubElement(rj, "parent", {"link": "world"})
    ET.SubElement(rj, "child", {"link": root})
    for j in src.findall("joint"):
        name, jt = j.get("name"), j.get("type")
        o = j.find("origin")
        xyz = [float(x) for x in (o.get("xyz", "0 0 0") if o is not None else "0 0 0").split()]
        rpy = [float(x) for x in (o.get("rpy", "0 0 0") if o is not None else "0 0 0").split()]
        if jt == "fixed":
            out.append(j)
            continue
        ax = j.find("axis")
        a = np.array([float(x) for x in (ax.get("xyz") if ax is not None else "1 0 0").split()])
        Ra = _align_x_to(a, rng)
        R = _rpy_mat(*rpy) @ Ra
        child = j.find("child").get("link")
        ej = ET.SubElement(out, "joint", {"name": name, "type": jt})
        ET.SubElement(ej, "origin", {"xyz": _fmt(xyz), "rpy": _fmt(_rpy_of(R))})
        ET.SubElement(ej, "axis", {"xyz": "1 0 0"})
        ET.SubElement(ej, "parent", {"link": j.find("parent").get("link")})
        ET.SubElement(ej, "child", {"link": f"{child}_dummy"})
        lim = j.find("limit")
        if lim is not None:
            ej.append(lim)
        ET.SubElement(out, "link", {"name": f"{child}_dummy"})
        dj = ET.SubElement(out, "joint", {"name": f"{name}_dummy", "type": "fixed"})
        ET.SubElement(dj, "origin", {"xyz": "0 0 0", "rpy": _fmt(_rpy_of(Ra.T))})
        ET.SubElement(dj, "axis", {"xyz": "0 0 0"})
        ET.SubElement(dj, "parent", {"link": f"{child}_dummy"})
        ET.SubElement(dj, "child", {"link": child})
    return ET.tostring(out, encoding="unicode")
