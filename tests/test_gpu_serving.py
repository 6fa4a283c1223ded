"""Continuous-batching serving loop (fem_hip_server_*).

* A request's solution is the same bit for bit alone and in a mix of concurrent
  requests (every solve uses the deterministic scatter and fixed reductions; the
  tuning table is not changed in between), both when small requests of one
  shape share cohort launches (batched, the default) and with one stream and
  graph per request, and equals the direct mass_solve of the same problem with
  the same scatter algorithm.
* Admission follows the HBM budget: with a budget that fits one request, only one
  runs at a time; a request larger than the budget fails with OUT_OF_MEMORY.
* Every request reaches its residual target; FP64 and mixed solves both do.
"""
import numpy as np
import pytest

import fehip
from fehip.problems import MassProblem
from fehip.serving import make_requests, run_workload

pytestmark = pytest.mark.gpu


def req_list():
    R = fehip.serve_request
    return [R(6, 3, 4, 1, fehip.SOLVE_FP64, mesh_seed=1, rhs_seed=11),
            R(5, 4, 5, 2, fehip.SOLVE_MIXED, mesh_seed=2, rhs_seed=12),
            R(9, 2, 3, 3, fehip.SOLVE_FP64, mesh_seed=3, rhs_seed=13),
            R(4, 6, 7, 1, fehip.SOLVE_MIXED, mesh_seed=4, rhs_seed=14),
            R(7, 5, 6, 1, fehip.SOLVE_FP64, mesh_seed=5, rhs_seed=15)]


@pytest.mark.parametrize("batched", [True, False])
def test_serve_alone_equals_in_mix(gpu, batched):
    """batched: the small requests share cohort launches; else one stream and
    graph per request."""
    print("serving: alone vs mix, batched", batched, flush=True)
    reqs = req_list()
    alone = []
    srv = fehip.Server(max_active=1, workers=2, hbm_fraction=0.3, batched=batched)
    for r in reqs:
        res, x = srv.wait(srv.submit(r), fehip.serve_l_size(r))
        assert res["status"] == 0 and res["rel_residual"] <= r.rtol * 1.0001, res
        alone.append(x)
    srv.close()
    # The same requests interleaved with 24 others, up to 12 at once.
    filler = make_requests(24, seed=5, n_range=(3, 10), p_range=(2, 6))
    mix = filler[:12] + reqs + filler[12:]
    summary, results = run_workload(mix, rate_per_s=400.0, keep_x=True, max_active=12, workers=6,
                                    hbm_fraction=0.3, batched=batched)
    assert summary["failed"] == 0, summary["errors"]
    assert summary["server"]["active_peak"] > 1
    for k, r in enumerate(reqs):
        res, x = results[12 + k]
        assert res["rel_residual"] <= r.rtol * 1.0001
        assert x.tobytes() == alone[k].tobytes(), k


def test_serve_matches_direct_solve(gpu):
    r = fehip.serve_request(6, 3, 4, 1, fehip.SOLVE_FP64, mesh_seed=0, rhs_seed=7)
    srv = fehip.Server(max_active=2, workers=1)
    res, x = srv.wait(srv.submit(r), fehip.serve_l_size(r))
    srv.close()
    p = MassProblem(6, 3, 4, 1, jitter=0.1, seed=0)
    # The server applies with the transpose-map algorithm when the shape is not in the tuning table.
    assert not [t for t in fehip.tune_table() if t[:3] == (3, 4, 1)]
    p.op.set_fused_algorithm(fehip.FUSED_ELEM_TMAP)
    b = fehip.DeviceArray.from_host(fehip.rng_uniform(7, p.l_size))
    d_x = fehip.DeviceArray(p.l_size)
    st = p.op.solve(b, d_x, mode=fehip.SOLVE_FP64, rtol=1e-10)
    assert st["iterations"] == res["iterations"]
    assert d_x.to_host().tobytes() == x.tobytes()


def test_serve_admission_by_budget(gpu):
    small = fehip.serve_request(6, 3, 4, 1, fehip.SOLVE_FP64)
    need = fehip.serve_request_bytes(small)
    info = fehip.device_info(0)
    frac = 1.5 * need / info.free_global_mem
    srv = fehip.Server(max_active=8, workers=4, hbm_fraction=frac)
    ids = [srv.submit(fehip.serve_request(6, 3, 4, 1, fehip.SOLVE_FP64, rhs_seed=s)) for s in range(6)]
    for i in ids:
        res, _ = srv.wait(i)
        assert res["status"] == 0
    st = srv.stats()
    assert st["active_peak"] == 1 and st["completed"] == 6 and st["used"] == 0
    assert st["used_peak"] <= st["budget"]
    big = srv.submit(fehip.serve_request(40, 6, 7, 3, fehip.SOLVE_FP64))
    with pytest.raises(fehip.FemHipError) as e:
        srv.wait(big)
    assert e.value.code == fehip.ERR_OUT_OF_MEMORY
    srv.close()


def test_serve_estimate_covers_actual_use(gpu):
    """The admission estimate is at least the device memory one request really takes."""
    r = fehip.serve_request(24, 5, 6, 2, fehip.SOLVE_MIXED)
    fehip.synchronize()
    before = fehip.device_info(0).free_global_mem
    srv = fehip.Server(max_active=1, workers=1, keep_x=False)
    rid = srv.submit(r)
    low = before
    while True:
        low = min(low, fehip.device_info(0).free_global_mem)
        got = srv.wait(rid, timeout_ms=1.0)
        if got is not None:
            break
    srv.close()
    assert got[0]["status"] == 0
    assert before - low <= fehip.serve_request_bytes(r), (before - low, fehip.serve_request_bytes(r))


def test_serve_invalid_request_fails_alone(gpu):
    srv = fehip.Server(max_active=2, workers=1)
    bad = srv.submit(fehip.serve_request(4, 7, 7, 1))  # no compiled (7, 7) kernel
    good = srv.submit(fehip.serve_request(4, 3, 4, 1))
    with pytest.raises(fehip.InvalidArgument):
        srv.wait(bad)
    assert srv.wait(good)[0]["status"] == 0
    srv.close()


def _residual(req, x):
    """||b - M x|| / ||b|| in FP64 on an independently built copy of the request's problem."""
    p = MassProblem(req.n, req.P, req.Q, req.num_comp, jitter=req.jitter, seed=req.mesh_seed)
    b = fehip.rng_uniform(req.rhs_seed, p.l_size)
    d_x, d_Mx = fehip.DeviceArray.from_host(x), fehip.DeviceArray(p.l_size)
    p.op.apply(d_x, d_Mx)
    return float(np.linalg.norm(b - d_Mx.to_host()) / np.linalg.norm(b))


def test_serve_workload_distribution_with_largest_shape(gpu):
    """Requests drawn from the serving workload's own distribution (n ~ U{4..40},
    P ~ U{2..6}, Q = P+1, num_comp ~ U{1..3}, FP64 or mixed), with the largest
    shape it can draw (n=40, P=6, 3 components: 24.4M DoF) added in both modes:
    every request reaches 1e-10 (the server's FP64 residual, and for the largest
    ones an FP64 residual recomputed here on an independently built problem),
    and a subset solved alone equals its result in the mix bit for bit."""
    print("serving: workload distribution with (40, 6, 3)", flush=True)
    R = fehip.serve_request
    big64 = R(40, 6, 7, 3, fehip.SOLVE_FP64, mesh_seed=77, rhs_seed=77)
    bigmx = R(40, 6, 7, 3, fehip.SOLVE_MIXED, mesh_seed=78, rhs_seed=78)
    drawn = make_requests(30, seed=11)
    mix = drawn[:15] + [big64] + drawn[15:] + [bigmx]
    summary, results = run_workload(mix, rate_per_s=300.0, keep_x=True, max_active=16, workers=8,
                                    hbm_fraction=0.5)
    print({k: summary[k] for k in ("completed", "failed", "throughput_req_per_s", "latency_ms")}, flush=True)
    assert summary["failed"] == 0, summary["errors"]
    assert summary["server"]["active_peak"] > 1
    modes = set()
    for r, (res, x) in zip(mix, results):
        assert res["status"] == 0 and res["rel_residual"] <= r.rtol, (r.n, r.P, r.num_comp, res)
        assert len(x) == res["l_size"]
        modes.add(r.mode)
    assert modes == {fehip.SOLVE_FP64, fehip.SOLVE_MIXED}
    for k in (15, len(mix) - 1):
        rel = _residual(mix[k], results[k][1])
        print("largest request", k, "recomputed residual", rel, flush=True)
        assert rel <= 1.01e-10, (k, rel)
    # Alone: the two largest and two drawn requests, one at a time.
    srv = fehip.Server(max_active=1, workers=2, hbm_fraction=0.5)
    for k in (0, 15, 16, len(mix) - 1):
        r = mix[k]
        res, x = srv.wait(srv.submit(r), fehip.serve_l_size(r))
        assert res["status"] == 0
        assert x.tobytes() == results[k][1].tobytes(), k
    srv.close()
