"""ctypes view of the C ABI (include/sapien_amd.h).

This is the binding a non-Python host (or a ctypes caller) uses: the same
``libsapien_amd.so`` the pybind module links against, called with raw device
pointers. ``launcher_fk_jac_rnea`` returns a minimal-overhead callable used by
the benchmark and the hipGraph capture tests.
"""
from __future__ import annotations

import ctypes
import os
import re

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libsapien_amd.so")
HEADER = os.path.join(os.path.dirname(_HERE), "include", "sapien_amd.h")
_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is not built; run `python sapien_amd/build.py`")
        L = ctypes.CDLL(LIB_PATH)
        P, I64, I32, D = ctypes.c_void_p, ctypes.c_int64, ctypes.c_int32, ctypes.c_double
        L.sapamd_last_error.restype = ctypes.c_char_p
        L.sapamd_version.restype = ctypes.c_char_p
        L.sapamd_fk_jac_rnea_batched.argtypes = [P, I64, P, P, P, P, P, P, P]
        L.sapamd_fk_jac_rnea_batched.restype = ctypes.c_int
        _lib = L
    return _lib


def declared_symbols():
    """Every function declared in include/sapien_amd.h."""
    text = open(HEADER).read()
    return sorted(set(re.findall(r"\b(sapamd_[a-z0-9_]+)\s*\(", text)))


def check(rc):
    if rc != 0:
        msg = lib().sapamd_last_error().decode()
        if rc == 1:
            raise ValueError(msg)
        raise RuntimeError(msg)


def handle(model):
    from . import _pinocchio

    return ctypes.c_void_p(_pinocchio._abi_handle(model))


def launcher_pose_torque_loss(model):
    """f(B, q, qd, qdd, tau_star, ee_link, p_target, loss, tau, dq, dqd, dqdd, stream) over tensors."""
    L = lib()
    h = handle(model)
    fn = L.sapamd_pose_torque_loss_batched
    P, I64, I32 = ctypes.c_void_p, ctypes.c_int64, ctypes.c_int32
    fn.argtypes = [P, I64, P, P, P, P, I32, P, P, P, P, P, P, P]
    fn.restype = ctypes.c_int

    def ptr(x):
        return None if x is None else (x if isinstance(x, int) else x.data_ptr())

    def launch(B, q, qd, qdd, tau_star, ee_link, p_target, loss, tau, dq, dqd, dqdd, stream):
        rc = fn(h, B, ptr(q), ptr(qd), ptr(qdd), ptr(tau_star), ee_link, ptr(p_target), ptr(loss), ptr(tau), ptr(dq),
                ptr(dqd), ptr(dqdd), stream)
        if rc:
            check(rc)

    return launch


def launcher_fk_jac_rnea(model):
    """f(B, q, qd, qdd, P, J, T, stream) over torch tensors (or raw ints)."""
    L = lib()
    h = handle(model)
    fn = L.sapamd_fk_jac_rnea_batched

    def ptr(x):
        if x is None:
            return None
        return x if isinstance(x, int) else x.data_ptr()

    def launch(B, q, qd, qdd, P, J, T, stream):
        rc = fn(h, B, ptr(q), ptr(qd), ptr(qdd), ptr(P), ptr(J), ptr(T), stream)
        if rc:
            check(rc)

    return launch


def launcher_fk_jac_rnea_timed(model):
    """f(B, q, qd, qdd, P, J, T, start_event, stop_event, stream): the fused
    launch with the kernel's own start and stop recorded into two HIP events
    (raw hipEvent_t handles, e.g. torch.cuda.Event(enable_timing=True).cuda_event
    after one record)."""
    L = lib()
    h = handle(model)
    fn = L.sapamd_fk_jac_rnea_batched_timed
    P, I64 = ctypes.c_void_p, ctypes.c_int64
    fn.argtypes = [P, I64, P, P, P, P, P, P, P, P, P]
    fn.restype = ctypes.c_int

    def ptr(x):
        if x is None:
            return None
        return x if isinstance(x, int) else x.data_ptr()

    def launch(B, q, qd, qdd, Pp, J, T, ev0, ev1, stream):
        rc = fn(h, B, ptr(q), ptr(qd), ptr(qdd), ptr(Pp), ptr(J), ptr(T), ev0, ev1, stream)
        if rc:
            check(rc)

    return launch


def launcher_hetero(models, B, q, qd, qdd, poses, J, tau, M):
    """f(stream) launching sapamd_hetero_batched (FK + J + RNEA + CRBA of
    several models in one launch) over fixed per-model device tensors (lists,
    one entry per model; an output entry may be None)."""
    L = lib()
    fn = L.sapamd_hetero_batched
    P, I32 = ctypes.c_void_p, ctypes.c_int32
    fn.argtypes = [P, I32, P, P, P, P, P, P, P, P, P]
    fn.restype = ctypes.c_int
    n = len(models)
    hs = (ctypes.c_void_p * n)(*[handle(m).value for m in models])
    Bs = (ctypes.c_int64 * n)(*B)

    def arr(xs):
        return (ctypes.c_void_p * n)(*[None if x is None else x.data_ptr() for x in xs])

    args = [hs, n, Bs, arr(q), arr(qd), arr(qdd), arr(poses), arr(J), arr(tau), arr(M)]

    def launch(stream):
        rc = fn(*args, ctypes.c_void_p(stream))
        if rc:
            check(rc)

    return launch
