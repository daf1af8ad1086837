"""Error metrics shared by the parity tests.

Tolerances of the batched FP32 path against the FP64 oracle:
  link pose      position |dp| <= 1e-5 m, rotation angle between quaternions <= 1e-5 rad
  Jacobian, tau  per environment max|d| / max|ref| <= 1e-4  (relative to the env's scale)
"""
import numpy as np

POSE_TOL_M = 1e-5
POSE_TOL_RAD = 1e-5
REL_TOL = 1e-4


def pose_errors(P, Pref):
    P = np.asarray(P, np.float64)
    pos = np.abs(P[..., :3] - Pref[..., :3]).max()
    qa = P[..., 3:] / np.linalg.norm(P[..., 3:], axis=-1, keepdims=True)
    dots = np.abs((qa * Pref[..., 3:]).sum(-1)).clip(0, 1)
    # angle = 2 asin(|q1 - (+-)q2| / 2) is better conditioned than acos near 0
    d = np.minimum(np.linalg.norm(qa - Pref[..., 3:], axis=-1), np.linalg.norm(qa + Pref[..., 3:], axis=-1))
    ang = (4 * np.arcsin(np.clip(d / 2, 0, 1))).max() if d.size else 0.0
    return pos, ang


def rel_err(X, Xref):
    """max over envs of max|X - Xref| / max|Xref| (env = leading axis)."""
    X = np.asarray(X, np.float64)
    B = X.shape[0]
    d = np.abs(X - Xref).reshape(B, -1).max(axis=1)
    s = np.abs(Xref).reshape(B, -1).max(axis=1)
    return float((d / np.maximum(s, 1e-12)).max()) if B else 0.0


def quat_to_R(qv):
    """Rotation matrices of (..., 4) quaternions w x y z (normalized first)."""
    w, x, y, z = np.moveaxis(qv / np.linalg.norm(qv, axis=-1, keepdims=True), -1, 0)
    return np.stack([np.stack([1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)], -1),
                     np.stack([2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)], -1),
                     np.stack([2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)], -1)], -2)


def link_coms(urdf, link_names):
    """Each link's own centre of mass in its frame (its URDF inertial origin)."""
    import xml.etree.ElementTree as ET

    out = []
    links = {l.get("name"): l for l in ET.fromstring(urdf).findall("link")}
    for n in link_names:
        o = links[n].find("inertial/origin")
        out.append([float(x) for x in o.get("xyz").split()] if o is not None else [0.0, 0.0, 0.0])
    return np.array(out)


def link_state_reference(o, urdf, link_names, q, qd, root):
    """World link poses and centre-of-mass / angular velocities of articulations
    placed by the pose and velocity of user link 0 (root rows p | q wxyz | v | w),
    from the oracle's model-frame poses and WORLD link Jacobians:
    (positions (B, L, 3), rotations (B, L, 3, 3), v_com (B, L, 3), w (B, L, 3))."""
    B, L = q.shape[0], len(link_names)
    Pm = o.link_poses(q)
    Rm, pm = quat_to_R(Pm[..., 3:]), Pm[..., :3]
    R0, p0, vc0, w0 = quat_to_R(root[:, 3:7]), root[:, :3], root[:, 7:10], root[:, 10:13]
    c = link_coms(urdf, link_names)
    Vm = np.stack([np.einsum("bij,bj->bi", o.link_jacobian(q, l, False), qd) for l in range(L)], 1)
    wm, vm = Vm[..., 3:], Vm[..., :3] + np.cross(Vm[..., 3:], pm)  # link origin velocity, model frame
    Rb = R0 @ np.swapaxes(Rm[:, 0], -1, -2)
    pb = p0 - np.einsum("bij,bj->bi", Rb, pm[:, 0])
    wb = w0 - np.einsum("bij,bj->bi", Rb, wm[:, 0])
    ub = vc0 - np.cross(w0, np.einsum("bij,j->bi", R0, c[0])) - np.einsum("bij,bj->bi", Rb, vm[:, 0])
    Rw = Rb[:, None] @ Rm
    pw = np.einsum("bij,blj->bli", Rb, pm) + pb[:, None]
    ww = wb[:, None] + np.einsum("bij,blj->bli", Rb, wm)
    vc = (ub[:, None] + np.cross(wb[:, None], pw - p0[:, None]) + np.einsum("bij,blj->bli", Rb, vm) +
          np.cross(ww, np.einsum("blij,lj->bli", Rw, c)))
    return pw, Rw, vc, ww
