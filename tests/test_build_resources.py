"""CPU: every product kernel stays within the private-segment (scratch) budget.

The compiler's resource remarks (-Rpass-analysis=kernel-resource-usage, kept
by the build beside each object) give each kernel's private segment per lane
and occupancy. Budgets (sapien_amd/build.py): lanes-per-joint kernels at most
4 KiB per lane, thread-per-environment kernels at most 64 KiB per lane, and
per kernel at most 16 GiB of backing store for a full device
(bytes per lane x 64 lanes x occupancy x 1024 SIMDs). A regression that spills
a lanes kernel's register arrays, or grows a per-joint array class, fails here
before it can exhaust the agent-wide scratch limit on the GPU.
"""
import importlib.util
import os

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _build_module():
    spec = importlib.util.spec_from_file_location("sapien_amd_build", os.path.join(ROOT, "sapien_amd", "build.py"))
    b = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(b)
    return b


def _stale(b):
    srcs = [os.path.join(b.CSRC, f) for f in os.listdir(b.CSRC)] + [os.path.join(ROOT, "include", "sapien_amd.h")]
    newest = max(os.path.getmtime(p) for p in srcs)
    for src in b.HIP_SOURCES:
        rep = os.path.join(b.OBJ, src + ".resources.txt")
        if not os.path.exists(rep) or os.path.getmtime(rep) < newest:
            return True
    return False


@pytest.fixture(scope="module")
def report():
    b = _build_module()
    if _stale(b):  # reports missing or older than the sources: rebuild the library
        b.build_lib(jobs=8)
    return b, b.resource_report()


def test_every_source_reports_kernels(report):
    b, rows = report
    assert {r["source"] for r in rows} >= {s for s in b.HIP_SOURCES if s.endswith(".hip")}
    assert all("private" in r and "occupancy" in r for r in rows)
    names = {b.kernel_base_name(r["name"]) for r in rows}
    for k in ("k_chain_fused", "k_tree_fused", "k_hetero", "k_clik_lanes", "k_tree_rollout", "k_tree_rollout_mixed",
              "k_kinematics", "k_aba", "k_rollout_step"):
        assert k in names, k


def test_private_segments_within_budget(report):
    b, rows = report
    bad = b.check_resources(rows)
    assert not bad, "\n".join(f"{n} ({s}): {p} B/lane, {r / 2**30:.1f} GiB" for n, s, p, r in bad)


def test_lanes_kernels_of_the_measured_paths_use_no_private_segment(report):
    """The headline chain kernel's shapes up to 8 joints per lane, the tree,
    mixed-tree, CLIK and rollout kernels: registers and LDS only."""
    b, rows = report
    for r in rows:
        base = b.kernel_base_name(r["name"])
        if base in ("k_tree_fused", "k_hetero", "k_clik_lanes", "k_tree_rollout", "k_tree_rollout_vjp",
                    "k_tree_rollout_mixed", "k_tree_step", "k_tree_step_vjp", "k_link_state"):
            assert r["private"] == 0, (r["name"], r["private"])


def test_budget_check_flags_an_oversized_kernel():
    b = _build_module()
    rows = [{"name": "_ZN6sapamd12_GLOBAL__N_113k_chain_fusedILi1ELi16EEEv", "source": "x", "private": 8192,
             "occupancy": 4},
            {"name": "_ZN6sapamd12_GLOBAL__N_15k_abaIdLi64ELb0EEEv", "source": "x", "private": 60000, "occupancy": 8},
            {"name": "_ZN6sapamd12_GLOBAL__N_15k_abaIdLi32ELb0EEEv", "source": "x", "private": 22032, "occupancy": 3}]
    flagged = [n for n, *_ in b.check_resources(rows)]
    assert flagged == [rows[0]["name"], rows[1]["name"]]
