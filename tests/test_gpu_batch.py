"""Batched mass solves (fem_hip_mass_solve_batched): solves of one shape and
precision share each persistent element and step launch of their cohort's
captured graph.

* Each batched solve reaches 1e-10 (its own FP64 residual, and one recomputed
  here with an independent FP64 apply) and the oracle CG's solution.
* A solve's result is the same bits alone and in a batch of other solves of
  the same shape and of other shapes (its arithmetic does not depend on who
  shares its launches), and across repeats.
* An FP64 batched solve equals the direct solver's pipelined PCG on the
  transpose-map algorithm bit for bit (the same element and step arithmetic),
  with the same iteration count.
* More solves of one shape than a cohort's table holds (16) wait for entries
  and all finish.
"""
import numpy as np
import pytest

import fehip
from fehip.problems import MassProblem
from helpers import rel_err
from oracle_lib import oracle

pytestmark = pytest.mark.gpu


def _rhs(p, seed):
    return fehip.DeviceArray.from_host(fehip.rng_uniform(seed, p.l_size, -1.0, 1.0))


def _true_residual(p, d_b, d_x):
    """||b - M x|| / ||b|| with an FP64 apply of the exact, deterministic algorithm."""
    p.op.set_fused_options(fehip.SCATTER_DETERMINISTIC, exact=1)
    d_mx = fehip.DeviceArray(p.l_size)
    p.op.apply(d_x, d_mx)
    p.op.set_fused_options(fehip.SCATTER_DETERMINISTIC, exact=0)
    b = d_b.to_host()
    return float(np.linalg.norm(b - d_mx.to_host()) / np.linalg.norm(b))


SHAPES = [(5, 3, 4, 1), (7, 3, 4, 1), (3, 3, 4, 1), (4, 5, 6, 2), (6, 2, 3, 3), (3, 6, 7, 1), (4, 4, 5, 3)]


@pytest.mark.parametrize("mode", [fehip.SOLVE_FP64, fehip.SOLVE_MIXED])
def test_batched_alone_equals_in_batch(gpu, mode):
    print("batched solves, mode", mode, flush=True)
    probs = [MassProblem(n, P, Q, nc, seed=k) for k, (n, P, Q, nc) in enumerate(SHAPES)]
    bs = [_rhs(p, 100 + k) for k, p in enumerate(probs)]
    xs = [fehip.DeviceArray(p.l_size) for p in probs]
    stats = fehip.mass_solve_batched([p.op for p in probs], bs, xs, modes=mode, rtol=1e-10)
    mix = [x.to_host() for x in xs]
    for k, (p, st) in enumerate(zip(probs, stats)):
        assert st["rel_residual"] <= 1e-10, (k, st)
        rel = _true_residual(p, bs[k], xs[k])
        assert rel <= 1.01e-10, (k, rel)
        if mode == fehip.SOLVE_MIXED:
            assert st["outer"] >= 2, st
    # Each alone (a cohort of one), twice.
    for k, p in enumerate(probs):
        x1 = fehip.DeviceArray(p.l_size)
        st1 = fehip.mass_solve_batched([p.op], [bs[k]], [x1], modes=mode, rtol=1e-10)[0]
        assert x1.to_host().tobytes() == mix[k].tobytes(), k
        assert st1["iterations"] == stats[k]["iterations"] and st1["outer"] == stats[k]["outer"]
    # A different mix: the first three (one cohort) in reverse order with two other shapes.
    order = [2, 1, 0, 5, 3]
    xs2 = [fehip.DeviceArray(probs[k].l_size) for k in order]
    fehip.mass_solve_batched([probs[k].op for k in order], [bs[k] for k in order], xs2, modes=mode, rtol=1e-10,
                             iters_per_launch=6)
    for k, x in zip(order, xs2):
        assert x.to_host().tobytes() == mix[k].tobytes(), k


def test_batched_fp64_equals_direct_pipelined_solve(gpu):
    """The batched FP64 solve runs the direct solver's pipelined step and element
    arithmetic: same iterations, same bits."""
    p = MassProblem(6, 3, 4, 1, seed=3)
    q = MassProblem(4, 3, 4, 1, seed=4)
    d_b, d_c = _rhs(p, 21), _rhs(q, 22)
    x_b, x_q = fehip.DeviceArray(p.l_size), fehip.DeviceArray(q.l_size)
    st = fehip.mass_solve_batched([p.op, q.op], [d_b, d_c], [x_b, x_q], modes=fehip.SOLVE_FP64, rtol=1e-10)
    d_x = fehip.DeviceArray(p.l_size)
    p.op.set_fused_algorithm(fehip.FUSED_ELEM_TMAP)
    direct = p.op.solve(d_b, d_x, mode=fehip.SOLVE_FP64, rtol=1e-10, cg_variant=fehip.CG_PIPELINED)
    assert direct["iterations"] == st[0]["iterations"], (direct, st[0])
    assert d_x.to_host().tobytes() == x_b.to_host().tobytes()


def test_batched_matches_oracle_cg(gpu):
    """FP64 and mixed batched solves reach the oracle CG's solution (the oracle
    CG runs on the restated operator_apply), with the oracle's iteration count
    within 2 in FP64."""
    p = MassProblem(8, 3, 4, 1, seed=0)
    orc = oracle().mass(3, p.num_elem, p.P, p.Q, p.nc, p.l_size, p.num_nodes, p.offsets, p.qdata())
    b = orc.apply(fehip.rng_uniform(2, p.l_size))
    x_cpu, it_cpu, rel_cpu = orc.cg(b, orc.diagonal(), rtol=1e-10)
    assert rel_cpu <= 1e-10
    d_b = fehip.DeviceArray.from_host(b)
    for mode in (fehip.SOLVE_FP64, fehip.SOLVE_MIXED):
        d_x = fehip.DeviceArray(p.l_size)
        st = fehip.mass_solve_batched([p.op], [d_b], [d_x], modes=mode, rtol=1e-10)[0]
        assert st["rel_residual"] <= 1e-10
        assert rel_err(d_x.to_host(), x_cpu) < 1e-8, mode
        if mode == fehip.SOLVE_FP64:
            assert abs(st["iterations"] - it_cpu) <= 2, (st, it_cpu)


def test_batched_more_solves_than_table_entries(gpu):
    """20 FP64 solves of one shape (a cohort's table holds 16) and two mixed
    ones: the rest wait for entries; all reach 1e-10, and two of them equal
    their solves alone."""
    probs = [MassProblem(3 + k % 4, 3, 4, 1, seed=k) for k in range(22)]
    bs = [_rhs(p, 300 + k) for k, p in enumerate(probs)]
    xs = [fehip.DeviceArray(p.l_size) for p in probs]
    modes = [fehip.SOLVE_FP64] * 20 + [fehip.SOLVE_MIXED] * 2
    stats = fehip.mass_solve_batched([p.op for p in probs], bs, xs, modes=modes, rtol=1e-10, iters_per_launch=4)
    for k, st in enumerate(stats):
        assert st["rel_residual"] <= 1e-10, (k, st)
    for k in (0, 19, 21):
        x1 = fehip.DeviceArray(probs[k].l_size)
        fehip.mass_solve_batched([probs[k].op], [bs[k]], [x1], modes=modes[k], rtol=1e-10)
        assert x1.to_host().tobytes() == xs[k].to_host().tobytes(), k


def test_batched_argument_errors(gpu):
    p = MassProblem(3, 3, 4, 1)
    d_b, d_x = _rhs(p, 1), fehip.DeviceArray(p.l_size)
    with pytest.raises(fehip.InvalidArgument):
        fehip.mass_solve_batched([p.op, p.op], [d_b, d_b], [d_x, d_x])
    with pytest.raises(fehip.InvalidArgument):
        fehip.mass_solve_batched([p.op], [d_b], [d_x], modes=fehip.SOLVE_MIXED_RELIABLE)
    assert fehip.mass_solve_batched([], [], []) == []
