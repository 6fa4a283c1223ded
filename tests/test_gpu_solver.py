"""Mass solves: FP64 Jacobi-PCG, FP32-inner / FP64-refined PCG and FP32 PCG with
FP64 reliable updates on the fused operator, checked against the oracle's FP64
CG on the same operator (the reference has no solver; the oracle CG runs on the
restated operator_apply)."""
import numpy as np
import pytest

import fehip
from fehip.problems import MassProblem
from helpers import rel_err
from oracle_lib import oracle

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("scatter", [fehip.SCATTER_ATOMIC, fehip.SCATTER_DETERMINISTIC])
def test_c1_solves(gpu, scatter):
    print("c1 solves, scatter", scatter, flush=True)
    p = MassProblem(16, 3, 4, 1, scatter=scatter)
    orc = oracle().mass(3, p.num_elem, 3, 4, 1, p.l_size, p.num_nodes, p.offsets, p.qdata())
    x_star = fehip.rng_uniform(2, p.l_size)
    b = orc.apply(x_star)
    diag = orc.diagonal()
    x_cpu, it_cpu, rel_cpu = orc.cg(b, diag, rtol=1e-10)
    assert rel_cpu <= 1e-10
    d_b, d_x = fehip.DeviceArray.from_host(b), fehip.DeviceArray(p.l_size)
    s64 = p.op.solve(d_b, d_x, mode=fehip.SOLVE_FP64, rtol=1e-10)
    x64 = d_x.to_host()
    assert s64["rel_residual"] <= 1e-10, s64
    assert abs(s64["iterations"] - it_cpu) <= 2, (s64, it_cpu)
    s_mx = p.op.solve(d_b, d_x, mode=fehip.SOLVE_MIXED, rtol=1e-10, inner_rtol=1e-4)
    xm = d_x.to_host()
    assert s_mx["rel_residual"] <= 1e-10, s_mx
    # The default inner tolerance (1e-5) reaches 1e-10 in two or three refinement
    # steps, with about the FP64 iteration count in all.
    s_def = p.op.solve(d_b, d_x, mode=fehip.SOLVE_MIXED, rtol=1e-10)
    assert s_def["rel_residual"] <= 1e-10 and s_def["outer"] <= 3, s_def
    assert s_def["iterations"] <= s64["iterations"] + 8, (s_def, s64)
    assert rel_err(d_x.to_host(), x_cpu) < 1e-8
    # Every mode agrees with the oracle's solution and the manufactured x* to solver accuracy.
    assert rel_err(x64, x_cpu) < 1e-8
    assert rel_err(xm, x_cpu) < 1e-8
    assert rel_err(xm, x_star) < 1e-8
    for inner in (1e-2, 1e-3, 1e-4):
        s_rel = p.op.solve(d_b, d_x, mode=fehip.SOLVE_MIXED_RELIABLE, rtol=1e-10, inner_rtol=inner)
        xr = d_x.to_host()
        assert s_rel["rel_residual"] <= 1e-10, (inner, s_rel)
        assert rel_err(xr, x_cpu) < 1e-8 and rel_err(xr, x_star) < 1e-8, inner
        # No restarts: with a replacement every 100x drop (the default) the
        # iteration count stays close to the FP64 one.
        if inner == 1e-2:
            assert s_rel["iterations"] <= s64["iterations"] + 4, (inner, s_rel, s64)
        print("reliable", inner, s_rel, flush=True)
    print("fp64", s64, "mixed", s_mx, "cpu iters", it_cpu, flush=True)


def test_solve_repeatable(gpu):
    p = MassProblem(8, 4, 5, 2, scatter=fehip.SCATTER_DETERMINISTIC)
    b = fehip.rng_uniform(5, p.l_size)
    d_b, d_x = fehip.DeviceArray.from_host(b), fehip.DeviceArray(p.l_size)
    p.op.set_fused_options(fehip.SCATTER_DETERMINISTIC)
    for mode in (fehip.SOLVE_MIXED, fehip.SOLVE_MIXED_RELIABLE):
        p.op.solve(d_b, d_x, mode=mode)
        x1 = d_x.to_host()
        p.op.solve(d_b, d_x, mode=mode)
        assert d_x.to_host().tobytes() == x1.tobytes(), mode


@pytest.mark.parametrize("n,P,Q,nc", [(16, 3, 4, 1), (8, 4, 5, 2)])
def test_solves_with_box_columns(gpu, n, P, Q, nc):
    """Solves on the box-column algorithm (the solver's apply on structured boxes
    of 1M element nodes and up): FP64 takes the oracle CG's iteration count
    (within 2), every mode reaches 1e-10 and the oracle's solution, and a second
    solve repeats the first bit for bit."""
    print("box-column solves", n, P, Q, nc, flush=True)
    p = MassProblem(n, P, Q, nc, scatter=fehip.SCATTER_DETERMINISTIC)
    p.op.set_fused_algorithm(fehip.FUSED_BOX_COLUMN)
    assert p.op.fused_choice()[0] == fehip.FUSED_BOX_COLUMN
    assert p.op.fused_choice(fp32=True)[0] == fehip.FUSED_BOX_COLUMN
    orc = oracle().mass(3, p.num_elem, P, Q, nc, p.l_size, p.num_nodes, p.offsets, p.qdata())
    b = orc.apply(fehip.rng_uniform(2, p.l_size))
    x_cpu, it_cpu, rel_cpu = orc.cg(b, orc.diagonal(), rtol=1e-10)
    d_b, d_x = fehip.DeviceArray.from_host(b), fehip.DeviceArray(p.l_size)
    for mode in (fehip.SOLVE_FP64, fehip.SOLVE_MIXED, fehip.SOLVE_MIXED_RELIABLE):
        st = p.op.solve(d_b, d_x, mode=mode, rtol=1e-10)
        x1 = d_x.to_host()
        assert st["rel_residual"] <= 1e-10, (mode, st)
        assert rel_err(x1, x_cpu) < 1e-8, mode
        if mode == fehip.SOLVE_FP64:
            assert abs(st["iterations"] - it_cpu) <= 2, (st, it_cpu)
        p.op.solve(d_b, d_x, mode=mode, rtol=1e-10)
        assert d_x.to_host().tobytes() == x1.tobytes(), mode


def _oracle_solution(p, seed=2):
    orc = oracle().mass(3, p.num_elem, p.P, p.Q, p.nc, p.l_size, p.num_nodes, p.offsets, p.qdata())
    b = orc.apply(fehip.rng_uniform(seed, p.l_size))
    x_cpu, it_cpu, rel_cpu = orc.cg(b, orc.diagonal(), rtol=1e-10)
    assert rel_cpu <= 1e-10
    return b, x_cpu, it_cpu


def test_solve_graphs_follow_option_changes_and_autotune(gpu):
    """A solve after an algorithm change, a smaller box layer count, an
    autotune or a tuning-table change (same operator, same b and x) runs the
    operator's current kernels: the graphs captured by the previous solve are
    not replayed (they may hold buffers that autotune or a reallocation freed),
    and every solve reaches the oracle's solution."""
    print("solve after option changes", flush=True)
    p = MassProblem(16, 3, 4, 1, scatter=fehip.SCATTER_DETERMINISTIC)
    b, x_cpu, _ = _oracle_solution(p)
    d_b, d_x = fehip.DeviceArray.from_host(b), fehip.DeviceArray(p.l_size)

    def solve(mode, want_alg=None):
        fp32 = mode != fehip.SOLVE_FP64
        if want_alg is not None:
            assert p.op.fused_choice(fp32)[0] == want_alg
        st = p.op.solve(d_b, d_x, mode=mode, rtol=1e-10)
        assert st["rel_residual"] <= 1e-10, (mode, want_alg, st)
        assert rel_err(d_x.to_host(), x_cpu) < 1e-8, (mode, want_alg)

    fehip.tune_table_clear()
    try:
        for mode in (fehip.SOLVE_FP64, fehip.SOLVE_MIXED, fehip.SOLVE_MIXED_RELIABLE):
            p.op.set_fused_algorithm(fehip.FUSED_BOX_COLUMN)
            solve(mode, fehip.FUSED_BOX_COLUMN)
            p.op.set_box_layers(2)  # more segments: the partial area grows (reallocated)
            solve(mode, fehip.FUSED_BOX_COLUMN)
            p.op.set_box_layers(0)
            p.op.set_fused_algorithm(fehip.FUSED_ELEM_TMAP)
            solve(mode, fehip.FUSED_ELEM_TMAP)
            p.op.set_fused_algorithm(fehip.FUSED_BLOCK_PARTIAL)
            solve(mode, fehip.FUSED_BLOCK_PARTIAL)
            p.op.set_fused_algorithm(fehip.FUSED_AUTO)
            (alg, ne, _), _ = p.op.autotune(fp32=mode != fehip.SOLVE_FP64, reps=2)  # frees unused plans
            assert p.op.fused_choice(mode != fehip.SOLVE_FP64) == (alg, ne)
            solve(mode, alg)
            # A table row set by hand changes the automatic choice with no option call.
            fp32 = mode != fehip.SOLVE_FP64
            other = fehip.FUSED_ELEM_TMAP if alg != fehip.FUSED_ELEM_TMAP else fehip.FUSED_BLOCK_PARTIAL
            p.op.set_fused_algorithm(other)
            ne_o = p.op.fused_choice(fp32)[1]  # its default packing
            p.op.set_fused_algorithm(fehip.FUSED_AUTO)
            box = int(fehip.box_detect(p.num_elem, 3, p.offsets) is not None)  # the row autotune wrote
            fehip.tune_table_set(3, 4, 1, int(fp32), 1, other, ne_o, 0.001, p.num_elem, box=box)
            solve(mode, other)
            fehip.tune_table_clear()
    finally:
        fehip.tune_table_clear()


def test_solve_iteration_limit_per_call(gpu):
    """max_iter belongs to each solve, not to the graphs the first solve
    captured: a short limit, then a long one, then a short one again, on the
    same operator, b and x."""
    print("solve iteration limits", flush=True)
    p = MassProblem(8, 3, 4, 1, scatter=fehip.SCATTER_DETERMINISTIC)
    d_b = fehip.DeviceArray.from_host(fehip.rng_uniform(5, p.l_size))
    d_x = fehip.DeviceArray(p.l_size)
    s1 = p.op.solve(d_b, d_x, mode=fehip.SOLVE_FP64, rtol=1e-10, max_iter=5)
    assert s1["iterations"] == 5 and s1["rel_residual"] > 1e-10, s1
    s2 = p.op.solve(d_b, d_x, mode=fehip.SOLVE_FP64, rtol=1e-10, max_iter=2000)
    assert s2["rel_residual"] <= 1e-10 and s2["iterations"] > 5, s2
    s3 = p.op.solve(d_b, d_x, mode=fehip.SOLVE_FP64, rtol=1e-10, max_iter=3)
    assert s3["iterations"] == 3, s3
    # Mixed: the limit is per inner FP32 solve; two outer steps of at most 3.
    m1 = p.op.solve(d_b, d_x, mode=fehip.SOLVE_MIXED, rtol=1e-10, max_iter=3, max_outer=2)
    assert m1["outer"] == 2 and m1["iterations"] <= 6 and m1["rel_residual"] > 1e-10, m1
    m2 = p.op.solve(d_b, d_x, mode=fehip.SOLVE_MIXED, rtol=1e-10)
    assert m2["rel_residual"] <= 1e-10, m2
    m3 = p.op.solve(d_b, d_x, mode=fehip.SOLVE_MIXED, rtol=1e-10, max_iter=2, max_outer=1)
    assert m3["outer"] == 1 and m3["iterations"] <= 2, m3


@pytest.mark.parametrize("alg", [fehip.FUSED_AUTO, fehip.FUSED_BOX_COLUMN])
def test_high_order_vector_mass_solves(gpu, alg):
    """The high-order vector mass (P=8, Q=10, three components) at a reduced
    mesh: FP64 PCG and the FP32-apply / FP64-refined solve both reach 1e-10 and
    the oracle CG's solution (the reference has no solver; the oracle CG runs on
    the restated operator_apply)."""
    print("P=8 Q=10 nc=3 solves, alg", alg, flush=True)
    p = MassProblem(4, 8, 10, 3, scatter=fehip.SCATTER_DETERMINISTIC)
    p.op.set_fused_algorithm(alg)
    b, x_cpu, it_cpu = _oracle_solution(p)
    d_b, d_x = fehip.DeviceArray.from_host(b), fehip.DeviceArray(p.l_size)
    s64 = p.op.solve(d_b, d_x, mode=fehip.SOLVE_FP64, rtol=1e-10)
    assert s64["rel_residual"] <= 1e-10 and abs(s64["iterations"] - it_cpu) <= 2, (s64, it_cpu)
    assert rel_err(d_x.to_host(), x_cpu) < 1e-8
    smx = p.op.solve(d_b, d_x, mode=fehip.SOLVE_MIXED, rtol=1e-10)
    assert smx["rel_residual"] <= 1e-10 and smx["outer"] >= 2, smx
    assert rel_err(d_x.to_host(), x_cpu) < 1e-8
    print("fp64", s64, "mixed", smx, "oracle iterations", it_cpu, flush=True)


@pytest.mark.parametrize("n,P,Q,nc", [(16, 3, 4, 1), (6, 5, 6, 3)])
def test_plain_and_pipelined_cg(gpu, n, P, Q, nc):
    """Every CG form (plain: five launches an iteration; pipelined, Ghysels &
    Vanroose 2014: three; plain with the wide kernels: eight) in FP64 and mixed
    precision reaches 1e-10 (true FP64 residual), takes the oracle CG's
    iteration count within 2 (FP64), agrees with the oracle's solution, and
    repeats bit for bit. Both sizes leave a tail of n % 2 (FP64) and n % 4
    (FP32) entries after the 16-byte packs."""
    print("plain vs pipelined CG", n, P, Q, nc, flush=True)
    p = MassProblem(n, P, Q, nc, scatter=fehip.SCATTER_DETERMINISTIC)
    b, x_cpu, it_cpu = _oracle_solution(p)
    d_b, d_x = fehip.DeviceArray.from_host(b), fehip.DeviceArray(p.l_size)
    assert p.l_size % 4 != 0
    for variant in (fehip.CG_PLAIN, fehip.CG_PIPELINED, fehip.CG_PLAIN_WIDE):
        for mode in (fehip.SOLVE_FP64, fehip.SOLVE_MIXED):
            st = p.op.solve(d_b, d_x, mode=mode, rtol=1e-10, cg_variant=variant)
            x1 = d_x.to_host()
            assert st["rel_residual"] <= 1e-10, (variant, mode, st)
            assert rel_err(x1, x_cpu) < 1e-8, (variant, mode)
            if mode == fehip.SOLVE_FP64:
                assert abs(st["iterations"] - it_cpu) <= 2, (variant, st, it_cpu)
            p.op.solve(d_b, d_x, mode=mode, rtol=1e-10, cg_variant=variant)
            assert d_x.to_host().tobytes() == x1.tobytes(), (variant, mode)
            print("variant", variant, "mode", mode, st, flush=True)
    # Residual replacement runs the plain form whatever is asked, wide if asked.
    st = p.op.solve(d_b, d_x, mode=fehip.SOLVE_MIXED_RELIABLE, rtol=1e-10, cg_variant=fehip.CG_PIPELINED)
    assert st["rel_residual"] <= 1e-10
    sw = p.op.solve(d_b, d_x, mode=fehip.SOLVE_MIXED_RELIABLE, rtol=1e-10, cg_variant=fehip.CG_PLAIN_WIDE)
    assert sw["rel_residual"] <= 1e-10 and rel_err(d_x.to_host(), x_cpu) < 1e-8, sw
    assert abs(sw["iterations"] - st["iterations"]) <= 2, (sw, st)


@pytest.mark.parametrize("n,P,Q,nc", [(16, 3, 4, 1), (6, 5, 6, 3)])
@pytest.mark.parametrize("alg", [fehip.FUSED_BLOCK_PARTIAL, fehip.FUSED_ELEM_TMAP, fehip.FUSED_AUTO])
def test_pipelined_split_pass_bit_equal(gpu, n, P, Q, nc, alg):
    """The pipelined form with the apply's second pass (block partial sums or the
    transpose map) folded into its step kernel gives the bits and iteration
    counts of the form that runs the pass as its own kernel, in FP64 and mixed
    precision, and the oracle's solution."""
    print("split second pass", n, P, Q, nc, alg, flush=True)
    p = MassProblem(n, P, Q, nc, scatter=fehip.SCATTER_DETERMINISTIC)
    p.op.set_fused_algorithm(alg)
    b, x_cpu, it_cpu = _oracle_solution(p)
    d_b, d_x = fehip.DeviceArray.from_host(b), fehip.DeviceArray(p.l_size)
    for mode in (fehip.SOLVE_FP64, fehip.SOLVE_MIXED):
        s1 = p.op.solve(d_b, d_x, mode=mode, rtol=1e-10, cg_variant=fehip.CG_PIPELINED)
        x1 = d_x.to_host()
        s2 = p.op.solve(d_b, d_x, mode=mode, rtol=1e-10, cg_variant=fehip.CG_PIPELINED, separate_second_pass=True)
        x2 = d_x.to_host()
        print("mode", mode, s1, s2, flush=True)
        assert s1["rel_residual"] <= 1e-10 and rel_err(x1, x_cpu) < 1e-8, (mode, s1)
        assert (s1["iterations"], s1["outer"]) == (s2["iterations"], s2["outer"]), (mode, s1, s2)
        assert np.array_equal(x1.view(np.uint64), x2.view(np.uint64)), mode
