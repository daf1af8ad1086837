"""GPU: the allocation-free CLIK entry point (sapamd_clik_batched_noalloc)
against sapamd_clik_batched, both called through the C ABI on the same inputs:
the same bits in every output, on two calls in a row (the second smaller, so
it runs over scratch the first call left behind), for both kernel families;
more problems than the scratch holds fail with a clear error.
"""
import ctypes

import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

from oracle.urdf_model import OracleModel  # noqa: E402
from sapien_amd import synthetic as S  # noqa: E402

DEV = "cuda:0"
MAX_PROBLEMS = 65536


def _fn(name):
    from sapien_amd.native import lib

    f = getattr(lib(), name)
    P, I64, I32, D = ctypes.c_void_p, ctypes.c_int64, ctypes.c_int32, ctypes.c_double
    f.argtypes = [P, I64, I32, P, P, P, D, I32, D, D, P, P, P, P, P]
    f.restype = ctypes.c_int
    return f


def _call(f, m, link, tgt, q0, nv):
    from sapien_amd.native import handle

    B = tgt.shape[0]
    q = torch.full((B, nv), float("nan"), device=DEV)
    ok = torch.full((B,), -7, dtype=torch.int32, device=DEV)
    err = torch.full((B, 6), float("nan"), device=DEV)
    it = torch.full((B,), -7, dtype=torch.int32, device=DEV)
    stream = torch.cuda.current_stream().cuda_stream
    rc = f(handle(m), B, link, tgt.data_ptr(), q0.data_ptr() if q0 is not None else None, None, 1e-4, 1000, 0.1,
           1e-6, q.data_ptr(), ok.data_ptr(), err.data_ptr(), it.data_ptr(), stream)
    return rc, (q, ok, err, it)


@pytest.mark.parametrize("variant", ["auto", "thread_per_env"])
def test_noalloc_clik_matches_clik_batched(variant):
    import sapien_amd as sa
    from sapien_amd.native import check

    rob = S.chain7()
    m = sa.PinocchioModel(rob.urdf, [0.0, 0.0, -9.81])
    m.set_joint_order(rob.joint_names)
    m.set_link_order(rob.link_names)
    sa.set_kernel_variant(m, variant)
    o = OracleModel(rob.urdf)
    o.set_joint_order(rob.joint_names)
    o.set_link_order(rob.link_names)
    ref_fn, na_fn = _fn("sapamd_clik_batched"), _fn("sapamd_clik_batched_noalloc")
    for B, seed in ((6000, 2200), (1500, 2201)):  # the second call smaller than the first
        qs = S.samples(rob, B, seed)[0]
        tgt = torch.tensor(o.link_poses(qs)[:, 7], dtype=torch.float32, device=DEV)
        tgt[::7, :3] += 4.0  # some unreachable targets: they run to the iteration limit
        q0 = torch.tensor(S.samples(rob, B, seed + 10)[0], dtype=torch.float32, device=DEV)
        rc, ref = _call(ref_fn, m, 7, tgt, q0, rob.nv)
        check(rc)
        rc, got = _call(na_fn, m, 7, tgt, q0, rob.nv)
        check(rc)
        torch.cuda.synchronize()
        for a, b in zip(got, ref):
            assert torch.equal(a, b), (variant, B)
        okn = ref[1].cpu().numpy()
        assert 0.5 < okn.mean() < 1.0 and not okn[::7].any()
    # beyond the scratch: a clear error, nothing launched
    big = torch.zeros(MAX_PROBLEMS + 1, 7, device=DEV)
    big[:, 3] = 1.0
    rc, _ = _call(na_fn, m, 7, big, None, rob.nv)
    assert rc == 1
    with pytest.raises(ValueError, match="at most 65536 problems"):
        check(rc)
