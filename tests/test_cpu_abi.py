"""CPU-side checks of the product library (no GPU work): it loads, exports every
entry point include/fem_hip.h declares, and its host-side setup code (quadrature,
Lagrange matrices, meshes, RNG, transpose map, element blocking) is pinned against
the oracle and against independent numpy statements."""
import ctypes
import os
import re

import numpy as np
import pytest

import fehip
from fehip.problems import host_box_offsets
from oracle_lib import oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    text = open(os.path.join(ROOT, "include", "fem_hip.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(fem_hip_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    L = fehip.lib()
    names = header_functions()
    assert len(names) > 50
    for n in names:
        assert hasattr(L, n), n
        assert n in fehip.SIGNATURES, "binding misses " + n
    assert set(fehip.SIGNATURES) <= set(names)
    assert b"gfx950" in fehip.lib().fem_hip_version()


def test_no_device_is_not_an_error_for_host_queries():
    n = fehip.device_count()
    assert n >= 0


@pytest.mark.parametrize("mode", [0, 1])
def test_host_basis_setup_bitwise_equals_oracle(mode):
    o = oracle()
    for P in range(2, 11):
        for Q in range(1, 13):
            if mode == 1 and Q < 2:
                continue
            ob = o.basis(3, 1, P, Q, mode)
            nodes = fehip.gauss_lobatto(P, weights=False)
            x, w = fehip.gauss_legendre(Q) if mode == 0 else fehip.gauss_lobatto(Q)
            I, G = fehip.lagrange_basis_matrix(nodes, x)
            for a, b in ((x, ob["q_ref"]), (w, ob["q_weight"]), (I, ob["interp"]), (G, ob["grad"])):
                assert a.tobytes() == b.tobytes(), (P, Q, mode)


def test_host_setup_rejects_bad_sizes():
    with pytest.raises(fehip.InvalidArgument):
        fehip.gauss_legendre(0)
    with pytest.raises(fehip.InvalidArgument):
        fehip.gauss_lobatto(1)


def splitmix64(seed, n):
    """Independent statement of splitmix64 -> (x >> 11) * 2^-53."""
    out, s, M = [], seed, (1 << 64) - 1
    for _ in range(n):
        s = (s + 0x9E3779B97F4A7C15) & M
        z = s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M
        z ^= z >> 31
        out.append((z >> 11) * 2.0 ** -53)
    return np.array(out)


def test_rng_matches_splitmix64():
    assert (fehip.rng_uniform(1, 1000, 0.0, 1.0) == splitmix64(1, 1000)).all()
    u = fehip.rng_uniform(7, 100000)
    assert u.min() >= -1 and u.max() < 1 and abs(u.mean()) < 0.01


def test_box_offsets_match_numpy():
    for args in ((3, 2, 2, 3, 1), (2, 2, 2, 5, 3), (4, 1, 1, 2, 1)):
        assert (fehip.box_offsets(*args) == host_box_offsets(*args)).all()


def test_box_coords_jitter_and_boundary():
    n, jit = 5, 0.1
    c = fehip.box_coords(n, jit, 0).reshape(3, n + 1, n + 1, n + 1)  # [axis][z][y][x]
    g = np.arange(n + 1) / n
    ref = np.stack(np.meshgrid(g, g, g, indexing="ij")[::-1])  # x, y, z grids in [z][y][x]
    d = c - ref
    interior = np.zeros((n + 1,) * 3, bool)
    interior[1:-1, 1:-1, 1:-1] = True
    assert (d[:, ~interior] == 0).all()
    assert np.abs(d[:, interior]).max() <= jit / n
    r = splitmix64(0, 3 * (n - 1) ** 3).reshape(-1, 3)
    np.testing.assert_allclose(d[0][interior], jit / n * (2 * r[:, 0] - 1), atol=1e-16)


def test_transpose_map_matches_oracle():
    off = fehip.box_offsets(3, 2, 2, 3, 1)
    l_size = 7 * 5 * 5
    for a, b in zip(fehip.transpose_map(12, 27, l_size, off), oracle().transpose_map(12, 27, l_size, off)):
        assert (a == b).all()


def test_block_plan_invariants():
    # The library checks its own plan (every element once, local indices map back).
    n, P = 6, 4
    N = (n * (P - 1) + 1) ** 3
    off = fehip.box_offsets(n, n, n, P, 1)
    st = fehip.block_plan_stats(n ** 3, P ** 3, N, off, 16)
    assert st["blocks"] == -(-n ** 3 // 16)
    assert st["interior_nodes"] + st["boundary_nodes"] == st["unique_nodes"]
    assert st["boundary_nodes"] < 0.5 * st["element_nodes"]
    # Independent count of block-boundary nodes cannot exceed the per-element-node atomics.
    perm = np.random.default_rng(1).permutation(n ** 3)
    st2 = fehip.block_plan_stats(n ** 3, P ** 3, N, off.reshape(n ** 3, -1)[perm].ravel(), 16)
    assert st2["boundary_nodes"] < 0.6 * st2["element_nodes"]
    # Structured hexes need at most 8 colours (2x2x2 parity); every L entry is written.
    assert st["max_colors"] <= 8 and st2["max_colors"] <= 8
    assert st["covers_all"] == 1 and st["rows"] == N - st["interior_nodes"]


def test_block_plan_node_cap_and_coverage():
    n, P = 5, 3
    N = (n * (P - 1) + 1) ** 3
    off = fehip.box_offsets(n, n, n, P, 1)
    free = fehip.block_plan_stats(n ** 3, P ** 3, N, off, 16)
    capped = fehip.block_plan_stats(n ** 3, P ** 3, N, off, 16, max_nodes=60)
    assert free["max_nodes"] > 60 >= capped["max_nodes"]
    assert capped["blocks"] > free["blocks"]
    # An unreferenced trailing L entry is not written by the block passes.
    assert fehip.block_plan_stats(n ** 3, P ** 3, N + 1, off, 16)["covers_all"] == 0
    # Interleaved components (node stride 2, comp_stride 1) still cover the L-vector.
    off2 = fehip.box_offsets(n, n, n, P, 2)
    st = fehip.block_plan_stats(n ** 3, P ** 3, 2 * N, off2, 8, num_comp=2, comp_stride=1)
    assert st["covers_all"] == 1
    # A single-element cap is allowed; below one element's nodes is refused.
    assert fehip.block_plan_stats(n ** 3, P ** 3, N, off, 16, max_nodes=P ** 3)["blocks"] == n ** 3
    with pytest.raises(fehip.InvalidArgument):
        fehip.block_plan_stats(n ** 3, P ** 3, N, off, 16, max_nodes=P ** 3 - 1)


def test_restriction_create_validation_before_any_device_work():
    with pytest.raises(fehip.InvalidArgument):
        fehip.elem_restriction_create(2, 2, 1, 1, 3, np.array([0, 1, 1, 3]))
    with pytest.raises(fehip.InvalidArgument):
        fehip.elem_restriction_create(0, 2, 1, 1, 3, np.array([], np.int32))
    with pytest.raises(fehip.InvalidArgument):
        fehip.elem_restriction_create(2, 2, 2, 0, 3, np.array([0, 1, 1, 2]))
    with pytest.raises(fehip.InvalidArgument):
        fehip.elem_restriction_create_strided(2, 3, 1, 5)


def test_bench_refuses_multi_gpu():
    import subprocess
    import sys

    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=60)
    assert r.returncode == 2 and "one GPU" in r.stderr


def test_serve_request_bytes_estimate():
    """The serving admission estimate grows with the mesh, the order and the components."""
    est = lambda *a: fehip.serve_request_bytes(fehip.serve_request(*a))
    assert est(8, 3) < est(16, 3) < est(16, 5) < est(16, 5, 6, 3)
    # At least the solver's FP64 work vectors and the qdata copies.
    n, P, Q = 20, 4, 5
    L = (n * (P - 1) + 1) ** 3
    assert est(n, P, Q) > 10 * 8 * L + 4 * 8 * n ** 3 * Q ** 3


def test_tune_table_save_load_roundtrip(tmp_path):
    fehip.tune_table_clear()
    try:
        fehip.tune_table_set(5, 6, 1, 0, 1, fehip.FUSED_BLOCK_PARTIAL, 14, 0.39, 262144)
        fehip.tune_table_set(5, 6, 1, 0, 1, fehip.FUSED_BOX_COLUMN, 16, 0.26, 262144, box=1)
        fehip.tune_table_set(3, 4, 2, 1, 0, fehip.FUSED_ELEM_ATOMIC, 16, 0.05, 4096)
        # One key, two mesh sizes: two rows; the same size again replaces its row.
        fehip.tune_table_set(5, 6, 1, 0, 1, fehip.FUSED_ELEM_TMAP, 7, 0.02, 4096, box=1)
        fehip.tune_table_set(5, 6, 1, 0, 1, fehip.FUSED_BLOCK_PARTIAL, 7, 0.03, 4096, box=1)
        assert sorted(r[6:8] + (r[9],) for r in fehip.tune_table() if r[:6] == (5, 6, 1, 0, 1, 1)) == [
            (fehip.FUSED_BLOCK_PARTIAL, 7, 4096), (fehip.FUSED_BOX_COLUMN, 16, 262144)]
        with pytest.raises(fehip.InvalidArgument):  # box columns only in a box row
            fehip.tune_table_set(5, 6, 1, 0, 1, fehip.FUSED_BOX_COLUMN, 16, 0.26, 262144, box=0)
        with pytest.raises(fehip.InvalidArgument):  # the size a row was measured on is required
            fehip.tune_table_set(5, 6, 1, 0, 1, fehip.FUSED_ELEM_TMAP, 7, 0.02, 0)
        path = str(tmp_path / "t.txt")
        fehip.tune_table_save(path)
        rows = sorted(fehip.tune_table())
        fehip.tune_table_clear()
        assert fehip.tune_table() == []
        with open(path, "a") as f:
            f.write("# comment\nbest 7 7 1 0 0 0 0 8 0.1 10\n")  # no (7, 7) kernel: skipped
            f.write("best 5 6 1 1 1 1 7 0.2 512\n")  # a row without the box field: box = 0
        assert fehip.tune_table_load(path) == 5
        got = sorted(fehip.tune_table())
        assert (5, 6, 1, 1, 1, 0, fehip.FUSED_ELEM_TMAP, 7) in [r[:8] for r in got]
        got = [r for r in got if r[:4] != (5, 6, 1, 1)]
        assert [r[:8] for r in got] == [r[:8] for r in rows]
        np.testing.assert_allclose([r[8] for r in got], [r[8] for r in rows], rtol=1e-5)
        with open(path, "a") as f:
            f.write("bogus 1 2 3\n")
        with pytest.raises(fehip.InvalidArgument):
            fehip.tune_table_load(path)
    finally:
        fehip.tune_table_clear()


REF_INCLUDE = "/root/reference/include"


@pytest.mark.skipif(not os.path.isdir(REF_INCLUDE), reason="the reference is mounted only in the development container")
def test_dropin_header_compiles_against_reference_headers(tmp_path):
    """fem_hip.hpp and the drop-in test source compile and link with the
    reference's own include/ and src/ in place of the oracle's (so the fem::hip
    types sit on the reference's fem:: types, not on a restatement of them)."""
    import subprocess

    out = tmp_path / "test_fem_hip_ref"
    srcs = [os.path.join("/root/reference/src", f) for f in
            ("basis.cpp", "elem-restriction.cpp", "tensor-contract.cpp", "qfunction.cpp", "operator.cpp")]
    cmd = ["g++", "-std=c++17", "-O0", "-ffp-contract=off", "-Wall", "-Werror", "-Wno-sign-compare",
           "-include", "limits", "-I" + REF_INCLUDE, "-I" + os.path.join(ROOT, "include"),
           "-I" + os.path.join(ROOT, "fe-demo-cpp_amd", "cpp"), "-o", str(out),
           os.path.join(ROOT, "tests", "cpp", "test_fem_hip.cpp")] + srcs + [
        "-L" + os.path.join(ROOT, "fe-demo-cpp_amd", "lib"), "-lfemhip"]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-4000:]
