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
        "sapien_amd native extension is not built; run `python -m sapien_amd.build` "
        f"(or __graft_entry__.build()) first: {e}") from e

from ._pinocchio import PinocchioModel, Pose  # noqa: E402

LIBRARY_PATH = os.path.join(_HERE, "libsapien_amd.so")


def native_fk_jac_rnea(model):
    """Minimal-overhead launcher of the fused batched entry point (C ABI via ctypes)."""
    from .native import launcher_fk_jac_rnea

    return launcher_fk_jac_rnea(model)


def kernel_variant(model):
    return "generic"


def set_kernel_variant(model, name):
    if name != "generic":
        raise ValueError(f"unknown kernel variant {name!r}")


__all__ = ["PinocchioModel", "Pose", "LIBRARY_PATH", "native_fk_jac_rnea", "kernel_variant", "set_kernel_variant"]
