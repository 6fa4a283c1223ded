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
