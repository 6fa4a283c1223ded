"""Serving workload: mass-solve requests with Poisson arrivals through the
continuous-batching server (fem_hip_server_*).

Each request is a jittered structured hex mesh with n ~ U{n_lo..n_hi} elements
per side, P ~ U{p_lo..p_hi}, Q = P + 1, num_comp ~ U{1..3}, an FP64 or a
mixed-precision (FP32-inner / FP64-refined) solve, and a right-hand side
b ~ U[-1,1] from its own seed. Arrivals are a Poisson process of the given rate:
a submitting thread sleeps until each arrival time, so requests arrive while
others are on the GPU and are admitted as slots and memory free up.
"""
import threading
import time

import numpy as np

from . import SOLVE_FP64, SOLVE_MIXED, Server, serve_l_size, serve_request


def make_requests(count, seed=0, n_range=(4, 40), p_range=(2, 6), mixed_share=0.5, rtol=1e-10, first_seed=1000):
    """Request r: its shape and mode from numpy's generator (seed), its vertex
    jitter and right-hand side from splitmix64 seed first_seed + r."""
    rng = np.random.default_rng(seed)
    reqs = []
    for i in range(count):
        n = int(rng.integers(n_range[0], n_range[1] + 1))
        P = int(rng.integers(p_range[0], p_range[1] + 1))
        nc = int(rng.integers(1, 4))
        mode = SOLVE_MIXED if rng.random() < mixed_share else SOLVE_FP64
        reqs.append(serve_request(n, P, P + 1, nc, mode, 0.1, rtol, mesh_seed=first_seed + i,
                                  rhs_seed=first_seed + i))
    return reqs


def measure_capacity(count=300, seed=7, **server_opts):
    """Service capacity (requests/s): `count` requests of the same distribution
    all submitted at once, completions over wall time."""
    summary, _ = run_workload(make_requests(count, seed=seed, first_seed=500_000), rate_per_s=1e6, **server_opts)
    return summary["throughput_req_per_s"], summary


def run_workload(reqs, rate_per_s, seed=1, server=None, keep_x=False, **server_opts):
    """Submits `reqs` at Poisson arrival times of `rate_per_s` and collects every
    result. Returns a summary dict (throughput, latency percentiles, server stats)
    and the per-request results (with solutions if keep_x)."""
    own = server is None
    srv = server or Server(keep_x=keep_x, **server_opts)
    gaps = np.random.default_rng(seed).exponential(1.0 / rate_per_s, len(reqs))
    arrivals = np.cumsum(gaps) - gaps[0]
    ids = [None] * len(reqs)
    submitted = threading.Event()

    def submitter():
        t0 = time.perf_counter()
        for i, r in enumerate(reqs):
            delay = arrivals[i] - (time.perf_counter() - t0)
            if delay > 0:
                time.sleep(delay)
            ids[i] = srv.submit(r)
        submitted.set()

    t_start = time.perf_counter()
    th = threading.Thread(target=submitter)
    th.start()
    results = [None] * len(reqs)
    errors = []
    for i, r in enumerate(reqs):
        while ids[i] is None:
            time.sleep(0.0005)
        try:
            results[i] = srv.wait(ids[i], serve_l_size(r) if keep_x else 0)
        except Exception as e:  # a failed request is reported, not fatal
            errors.append((i, str(e)))
    th.join()
    wall = time.perf_counter() - t_start
    stats = srv.stats()
    if own:
        srv.close()
    ok = [res for res in results if res is not None]
    lat = np.array([res[0]["latency_ms"] for res in ok]) if ok else np.zeros(1)
    dofs = sum(res[0]["l_size"] for res in ok)
    iters = sum(res[0]["iterations"] for res in ok)
    summary = {
        "requests": len(reqs), "completed": len(ok), "failed": len(errors), "errors": errors[:5],
        "wall_s": wall, "offered_rate_per_s": rate_per_s, "throughput_req_per_s": len(ok) / wall,
        "solved_dof_per_s": dofs / wall, "cg_iterations_per_s": iters / wall,
        "latency_ms": {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99)),
                       "mean": float(lat.mean()), "max": float(lat.max())},
        "queue_ms_p50": float(np.median([res[0]["queue_ms"] for res in ok])) if ok else 0.0,
        "setup_ms_p50": float(np.median([res[0]["setup_ms"] for res in ok])) if ok else 0.0,
        "setup_build_ms_p50": float(np.median([res[0]["build_ms"] for res in ok])) if ok else 0.0,
        "setup_graph_ms_p50": float(np.median([res[0]["graph_ms"] for res in ok])) if ok else 0.0,
        "solve_ms_p50": float(np.median([res[0]["solve_ms"] for res in ok])) if ok else 0.0,
        "max_rel_residual": float(max(res[0]["rel_residual"] for res in ok)) if ok else None,
        "server": stats,
    }
    return summary, results
