"""sapien_amd — MI355X-native articulated-body kinematics and dynamics.

Drop-in for SAPIEN's ``PinocchioModel`` (python/py_package/wrapper/pinocchio_model.py,
pinocchio/pinocchio_model.cpp): the same class, method and argument names, with
every computation running in hand-written HIP kernels for gfx950 behind the C
ABI of ``include/sapien_amd.h``. Batched ``*_batched`` methods take (B, dof)
tensors through DLPack.

The native extension is required: importing this package without it raises.
"""
import os

_HERE = os.path.dirname(os.path.abspath(__file__))

try:
    from . import _pinocchio  # noqa: F401
except ImportError as e:  # pragma: no cover - exercised only on broken installs
    raise ImportError(
        "sapien_amd native extension is not built or is stale; run `python sapien_amd/build.py` "
        f"(or __graft_entry__.build()) first: {e}") from e

from ._pinocchio import (PinocchioModel, Pose, clik_hetero_batched, hetero_batched,  # noqa: E402
                         rollout_backward_hetero_batched, rollout_hetero_batched)

LIBRARY_PATH = os.path.join(_HERE, "libsapien_amd.so")


def native_fk_jac_rnea(model):
    """Minimal-overhead launcher of the fused batched entry point (C ABI via ctypes)."""
    from .native import launcher_fk_jac_rnea

    return launcher_fk_jac_rnea(model)


def native_fk_jac_rnea_timed(model):
    """The same launcher recording the kernel's own start and stop into two HIP
    events (kernel timing without a profiler)."""
    from .native import launcher_fk_jac_rnea_timed

    return launcher_fk_jac_rnea_timed(model)


_VARIANTS = {"auto": 0, "generic": 1, "thread_per_env": 1, "fast": 2, "lanes_per_joint": 2}
_VARIANT_NAMES = {1: "thread_per_env", 2: "lanes_per_joint"}


def kernel_variant(model):
    """Kernel family the batched kinematics methods launch for this model."""
    import ctypes

    from .native import check, handle, lib

    v = ctypes.c_int32(0)
    check(lib().sapamd_get_kernel_variant(handle(model), ctypes.byref(v)))
    return _VARIANT_NAMES[v.value]


def set_kernel_variant(model, name):
    """Select 'auto', 'thread_per_env' (alias 'generic') or 'lanes_per_joint' (alias 'fast')."""
    from .native import check, handle, lib

    if name not in _VARIANTS:
        raise ValueError(f"unknown kernel variant {name!r}")
    check(lib().sapamd_set_kernel_variant(handle(model), _VARIANTS[name]))


_ROLLOUT_LAUNCH = {"fused": 0, "per_step": 1}


def set_rollout_launch(model, mode):
    """'fused' (default): one launch per rollout (and per adjoint) for trees of
    <= 32 joints under the lanes-per-joint kernels; 'per_step': one launch per step."""
    from .native import check, handle, lib

    if mode not in _ROLLOUT_LAUNCH:
        raise ValueError(f"unknown rollout launch mode {mode!r}")
    check(lib().sapamd_set_rollout_launch(handle(model), _ROLLOUT_LAUNCH[mode]))


_PRECISIONS = {"auto": 0, "fp32": 1, "fp64": 2}


def dynamics_precision(model):
    """('fp32' | 'fp64', cond_estimate): the arithmetic of batched forward
    dynamics and rollouts for this model, and the largest cond(M) seen over
    64 seeded configurations within the joint limits."""
    import ctypes

    from .native import check, handle, lib

    mode, cond = ctypes.c_int32(0), ctypes.c_double(0)
    check(lib().sapamd_get_dynamics_precision(handle(model), ctypes.byref(mode), ctypes.byref(cond)))
    return ("fp64" if mode.value == 2 else "fp32"), cond.value


def set_dynamics_precision(model, name):
    """'auto' (FP64 factorization when cond(M) > 1e4), 'fp32' (ABA) or 'fp64'."""
    from .native import check, handle, lib

    if name not in _PRECISIONS:
        raise ValueError(f"unknown dynamics precision {name!r}")
    check(lib().sapamd_set_dynamics_precision(handle(model), _PRECISIONS[name]))


def _ik_multistart(self, link_index, poses, initial_qpos=None, n_init_qpos=20, active_qmask=None, eps=1e-4,
                   max_iterations=1000, dt=0.1, damp=1e-6, seed=0, stream=None):
    """Multi-start batched IK: see sapien_amd.ik.inverse_kinematics_multistart."""
    from .ik import inverse_kinematics_multistart

    return inverse_kinematics_multistart(self, link_index, poses, initial_qpos, n_init_qpos, active_qmask, eps,
                                         max_iterations, dt, damp, seed, stream)


PinocchioModel.compute_inverse_kinematics_multistart_batched = _ik_multistart

__all__ = ["PinocchioModel", "Pose", "hetero_batched", "clik_hetero_batched", "rollout_hetero_batched",
           "rollout_backward_hetero_batched", "LIBRARY_PATH", "native_fk_jac_rnea", "kernel_variant",
           "set_kernel_variant", "dynamics_precision", "set_dynamics_precision", "set_rollout_launch"]
