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
