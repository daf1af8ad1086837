"""Brute-force mechanics straight from URDF text, used to pin the oracle.

Nothing here shares code with the oracle or the product: the URDF is walked
link by link with 4x4 homogeneous matrices (no fixed-joint merging, no joint
tables, no spatial algebra), and the dynamics come from the Lagrangian of the
individual URDF links:

* link poses: T_child = T_parent @ origin @ joint_motion(q);
* per-link COM linear Jacobian and angular Jacobian from the geometric
  definition (revolute: a x (c - o), a; prismatic: a, 0);
* M(q) = sum_l m J_v^T J_v + J_w^T (R I R^T) J_w;
* V(q) = -sum_l m g . c_l;  g(q) = dV/dq = -sum_l m J_v^T g;
* Coriolis/centrifugal b(q, qd) = Mdot qd - 1/2 d(qd^T M qd)/dq by central
  differences of M.
"""
import math
import xml.etree.ElementTree as ET

import numpy as np


def _rpy(r, p, y):
    cr, sr, cp, sp, cy, sy = math.cos(r), math.sin(r), math.cos(p), math.sin(p), math.cos(y), math.sin(y)
    return np.array([[cy, -sy, 0], [sy, cy, 0], [0, 0, 1]]) @ np.array([[cp, 0, sp], [0, 1, 0], [-sp, 0, cp]]) @ \
        np.array([[1, 0, 0], [0, cr, -sr], [0, sr, cr]])


def _T(elem):
    T = np.eye(4)
    if elem is not None:
        xyz = [float(x) for x in elem.get("xyz", "0 0 0").split()]
        rpy = [float(x) for x in elem.get("rpy", "0 0 0").split()]
        T[:3, :3] = _rpy(*rpy)
        T[:3, 3] = xyz
    return T


def _rot(a, t):
    a = a / np.linalg.norm(a)
    K = np.array([[0, -a[2], a[1]], [a[2], 0, -a[0]], [-a[1], a[0], 0]])
    return np.eye(3) + math.sin(t) * K + (1 - math.cos(t)) * K @ K


class BruteRobot:
    def __init__(self, urdf, joint_names):
        root = ET.fromstring(urdf)
        self.links = {}
        for e in root.findall("link"):
            ie = e.find("inertial")
            if ie is None:
                self.links[e.get("name")] = (0.0, np.eye(4), np.zeros((3, 3)))
                continue
            m = float(ie.find("mass").get("value"))
            ii = ie.find("inertia")
            g = lambda k: float(ii.get(k, "0"))
            I = np.array([[g("ixx"), g("ixy"), g("ixz")], [g("ixy"), g("iyy"), g("iyz")], [g("ixz"), g("iyz"), g("izz")]])
            self.links[e.get("name")] = (m, _T(ie.find("origin")), I)
        self.joints = []
        child_of = {}
        for e in root.findall("joint"):
            ax = e.find("axis")
            a = np.array([float(x) for x in (ax.get("xyz") if ax is not None else "1 0 0").split()])
            j = dict(name=e.get("name"), type=e.get("type"), parent=e.find("parent").get("link"),
                     child=e.find("child").get("link"), T=_T(e.find("origin")), axis=a)
            self.joints.append(j)
            child_of[j["child"]] = j
        self.child_of = child_of
        self.root = [l for l in self.links if l not in child_of][0]
        self.dof = {n: i for i, n in enumerate(joint_names)}
        self.n = len(joint_names)
        # order links root-first
        order, frontier = [self.root], [self.root]
        while frontier:
            nxt = []
            for l in frontier:
                for j in self.joints:
                    if j["parent"] == l:
                        order.append(j["child"])
                        nxt.append(j["child"])
            frontier = nxt
        self.order = order

    def transforms(self, q):
        """World 4x4 of every URDF link plus per-dof world joint origin/axis."""
        T = {self.root: np.eye(4)}
        axes = {}
        for l in self.order[1:]:
            j = self.child_of[l]
            Tj = T[j["parent"]] @ j["T"]
            a = j["axis"] / np.linalg.norm(j["axis"])
            M = np.eye(4)
            if j["type"] in ("revolute", "continuous"):
                M[:3, :3] = _rot(a, q[self.dof[j["name"]]])
            elif j["type"] == "prismatic":
                M[:3, 3] = a * q[self.dof[j["name"]]]
            if j["name"] in self.dof:
                axes[j["name"]] = (Tj[:3, 3].copy(), Tj[:3, :3] @ a, j["type"])
            T[l] = Tj @ M
        return T, axes

    def _path_dofs(self, link):
        out = []
        l = link
        while l in self.child_of:
            j = self.child_of[l]
            if j["name"] in self.dof:
                out.append(j["name"])
            l = j["parent"]
        return out

    def link_jacobians(self, q):
        T, axes = self.transforms(q)
        res = {}
        for l, (m, To, I) in self.links.items():
            Tc = T[l] @ To
            c = Tc[:3, 3]
            Jv = np.zeros((3, self.n))
            Jw = np.zeros((3, self.n))
            for jn in self._path_dofs(l):
                o, a, t = axes[jn]
                k = self.dof[jn]
                if t == "prismatic":
                    Jv[:, k] = a
                else:
                    Jv[:, k] = np.cross(a, c - o)
                    Jw[:, k] = a
            Rl = T[l][:3, :3] @ To[:3, :3]
            res[l] = (m, c, Rl @ I @ Rl.T, Jv, Jw)
        return res

    def mass_matrix(self, q):
        M = np.zeros((self.n, self.n))
        for m, c, Iw, Jv, Jw in self.link_jacobians(q).values():
            M += m * Jv.T @ Jv + Jw.T @ Iw @ Jw
        return M

    def gravity_torque(self, q, g=(0, 0, -9.81)):
        g = np.asarray(g, float)
        out = np.zeros(self.n)
        for m, c, Iw, Jv, Jw in self.link_jacobians(q).values():
            out -= m * Jv.T @ g
        return out

    def potential(self, q, g=(0, 0, -9.81)):
        return -sum(m * np.dot(g, c) for m, c, *_ in self.link_jacobians(q).values())

    def bias(self, q, qd, h=1e-6):
        """b = Mdot qd - 1/2 d(qd^T M qd)/dq, central differences of M."""
        n = self.n
        Mdot = (self.mass_matrix(q + h * qd) - self.mass_matrix(q - h * qd)) / (2 * h)
        grad = np.zeros(n)
        for k in range(n):
            e = np.zeros(n)
            e[k] = h
            grad[k] = (qd @ self.mass_matrix(q + e) @ qd - qd @ self.mass_matrix(q - e) @ qd) / (2 * h)
        return Mdot @ qd - 0.5 * grad

    def inverse_dynamics(self, q, qd, qdd, g=(0, 0, -9.81)):
        return self.mass_matrix(q) @ qdd + self.bias(q, qd) + self.gravity_torque(q, g)
