#!/usr/bin/env python3
"""Mass-operator benchmark on one MI355X.

Workload (the headline configuration): FP64 fused mass operator_apply on a 64^3
hex mesh of the unit cube with +-0.1h vertex jitter (seed 0), P=5 GLL nodes, Q=6
Gauss points, num_comp=1: 257^3 = 16,974,593 DoF, u ~ U[-1,1] (seed 1). A step is
one operator_apply (every kernel of it: output zeroing or the deterministic
transpose-map pass, and the fused gather/interp/qdata/interp^T/scatter kernel)
with u, v, qdata and offsets resident in HBM.

Printed (one JSON line): GDoF/s of the fastest FP64 apply variant, the fused
kernel's HBM roofline (its algorithmic bytes per launch / its HIP-event duration),
every other apply variant (atomic, block-accumulated, exact, FP32), the 48^3,
P=8, Q=10, 3-component apply, the mixed-precision vs FP64 mass solves on the
16^3, P=3, Q=4 mesh (p50/p99 latency), the serving workload (2000 mass-solve
requests with Poisson arrivals through the continuous-batching server:
throughput and p50/p99 latency), and the CPU oracle timed on a bounded slab of
the headline workload.

--variants and --calibrate-copy serve profiler runs: the first restricts the
apply variants, the second adds K plain copies of one L-vector (a backend-strided
restriction gather: 8 bytes read and 8 written per DoF) so that counter readings
can be scaled by a kernel of known traffic.
"""
import argparse
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "fe-demo-cpp_amd"))

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, published


def log(*a):
    print(*a, file=sys.stderr, flush=True)


# The GPU's clocks settle only after some load: on MI355X the first
# milliseconds after an idle spell ran up to 20% slower. Every timed region is
# therefore preceded by at least this much back-to-back work, whatever --warmup
# asks for.
WARM_S = 0.25


def warm_up(op, d_u, d_v, warmup, stream, seconds=WARM_S):
    """`warmup` applies, then more until `seconds` of wall time have passed."""
    for _ in range(warmup):
        op.apply(d_u, d_v, stream)
    stream.synchronize()
    t0 = time.perf_counter()
    n = 0
    while time.perf_counter() - t0 < seconds:
        for _ in range(20):
            op.apply(d_u, d_v, stream)
        stream.synchronize()
        n += 20
    return n


def time_applies(fehip, op, d_u, d_v, steps, warmup, stream):
    """(ms per apply, ms in the main kernel, ms in the other stages). The apply
    time is the product path's: `steps` back-to-back applies between two
    events, after warm_up. The split comes from a second run of `steps`
    applies with the operator's stage events on (four events per apply, which
    add their own gaps, so that run is not the apply time)."""
    warm_up(op, d_u, d_v, warmup, stream)
    e0, e1 = fehip.Event(), fehip.Event()
    e0.record(stream)
    for _ in range(steps):
        op.apply(d_u, d_v, stream)
    e1.record(stream)
    e1.synchronize()
    total_ms = e0.elapsed_ms(e1)
    op.timing_begin()
    for _ in range(steps):
        op.apply(d_u, d_v, stream)
    stream.synchronize()
    k_ms, aux_ms, n = op.timing_end()
    return total_ms / steps, k_ms / max(n, 1), aux_ms / max(n, 1)


def copy_ceiling(fehip, stream, n=1 << 27, reps=20):
    """The device's streaming bandwidth: an n-double (1 GiB) array copied by the
    streaming copy kernel (one 16-byte nontemporal load and store per thread),
    each copy timed alone with HIP events on `stream`; the median's read plus
    written bytes per second."""
    a, b = fehip.DeviceArray(n), fehip.DeviceArray(n)
    a.zero(stream.h)
    b.zero(stream.h)
    for _ in range(5):
        fehip.stream_copy(b, a, stream)
    e0, e1 = fehip.Event(), fehip.Event()
    times = []
    for _ in range(reps):
        e0.record(stream)
        fehip.stream_copy(b, a, stream)
        e1.record(stream)
        e1.synchronize()
        times.append(e0.elapsed_ms(e1))
    ms = float(np.median(times))
    return {"GBs": 16.0 * n / (ms * 1e-3) / 1e9, "bytes_moved": 16 * n, "ms": ms}


def traffic_from_profile(kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 PMC summary
    (FETCH_SIZE + WRITE_SIZE, scaled by the copy-kernel calibration), if it holds
    that kernel."""
    path = os.path.join(ROOT, "profiles", "fused_mass_p5q6_64cube_pmc.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f).get("traffic_bytes_per_launch", {}).get(kernel)


def cpu_baseline(problem, v_gpu=None, slab_layers=8, reps=3):
    """The oracle's operator_apply (CPU restatement of the reference, one core) on a
    z-slab of the same mesh: the first `slab_layers` element layers of 64^3, on
    the slab's part of the same u (splitmix64 seed 1). With v_gpu (the device
    apply of the whole mesh) also the device's error against it on the slab's
    interior rows."""
    import fehip

    p = problem
    op, take, l_s, rows = slab_oracle(p, slab_layers)
    u = fehip.rng_uniform(1, l_s)
    # One core, as the reference runs (single-threaded): pin to the first allowed CPU.
    allowed = os.sched_getaffinity(0)
    os.sched_setaffinity(0, {min(allowed)})
    try:
        v = op.apply(u)
        times = []
        for _ in range(reps):
            t0 = time.perf_counter()
            op.apply(u)
            times.append(time.perf_counter() - t0)
    finally:
        os.sched_setaffinity(0, allowed)
    t = float(np.median(times))
    out = {"value": l_s / t / 1e9, "unit": "GDoF/s", "cores": 1, "kind": "port",
           "sample": "oracle operator_apply (single-threaded CPU restatement of the reference, -O2 "
                     "-ffp-contract=off) on a 64x64x%d-element z-slab of the 64^3 P=5 Q=6 mesh (%d DoF), "
                     "median of %d applies pinned to one core: %.3f s/apply; host %s, nproc %d" %
                     (slab_layers, l_s, reps, t, platform.processor() or platform.machine(), os.cpu_count())}
    if v_gpu is not None:
        out["gpu_rel_err_on_slab"] = slab_rel_err(v_gpu, v, l_s, rows, 1, p.num_nodes)
    return out


def slab_oracle(p, layers):
    """The oracle's mass operator on the first `layers` element layers (a z-slab)
    of problem p, with the qdata p built on the device. Returns (op, take, l_s,
    rows): take(vec) cuts a full L-vector to the slab's l_s nodes per component
    (every component), rows is the number of leading nodes per component whose
    rows the slab's elements alone make (every node below the slab's top node
    plane)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from oracle_lib import oracle

    n, P, Q, nc = p.n, p.P, p.Q, p.nc
    E_s = n * n * layers
    N1 = n * (P - 1) + 1
    l_s = N1 * N1 * (layers * (P - 1) + 1)
    off = p.offsets[: E_s * P ** 3]
    qd = p.qdata().reshape(Q ** 3, p.num_elem)[:, :E_s].copy().ravel()
    op = oracle().mass(3, E_s, P, Q, nc, nc * l_s, l_s, off, qd)
    N = p.num_nodes

    def take(vec):
        return np.concatenate([vec[c * N: c * N + l_s] for c in range(nc)])

    return op, take, l_s, N1 * N1 * layers * (P - 1)


def slab_rel_err(gpu_vec, cpu_slab, l_s, rows, nc, N):
    """max |gpu - cpu| / max |cpu| over the slab-interior rows of every component."""
    g = np.concatenate([gpu_vec[c * N: c * N + rows] for c in range(nc)])
    c_ = np.concatenate([cpu_slab[c * l_s: c * l_s + rows] for c in range(nc)])
    return float(np.max(np.abs(g - c_)) / np.max(np.abs(c_)))


def box_kernel_name(P, Q, nc, fp32, ne):
    """The demangled name of the box-column kernel a packing (elems) runs: the
    cross-section (elems % 100) and the variant (elems // 100: 0 qdata a layer
    ahead, 1 qdata in its own layer, 2 FP64 through the LDS ring, 3 the
    component split, 4 / 6 one / two elements per wavefront, 5 / 7 the same
    with the component split)."""
    bx, by = {16: (4, 4), 8: (4, 2), 4: (2, 2), 18: (8, 2), 32: (8, 4), 34: (16, 2)}[ne % 100]
    t, code = "float" if fp32 else "double", ne // 100
    if code >= 4:
        split = code in (5, 7)
        return "fused_box_wave_kernel<%d, %d, %d, %s, %d, %d, %d, %d, 0>" % (
            P, Q, 1 if split else nc, t, bx, by, nc if split else 1, 1 if code < 6 else 2)
    um = 2 if fp32 or code == 2 else 0
    if code == 3:
        return "fused_box_kernel<%d, %d, 1, %s, %d, %d, %d, 0, %d>" % (P, Q, t, bx, by, um, nc)
    return "fused_box_kernel<%d, %d, %d, %s, %d, %d, %d, %d, 1>" % (P, Q, nc, t, bx, by, um, 0 if code == 1 else 1)


def config5(fehip, args, stream, layers=2):
    """48^3 hexes, P=8, Q=10, 3 components (114.8M DoF): the FP64 and FP32 applies
    of each algorithm (GDoF/s, %HBM), the FP64 apply's error against the oracle's
    operator_apply on a 2-layer z-slab (its interior rows), and the FP64 and
    FP32-apply / FP64-refined solves of M x = b, b = M x* (x* ~ U[-1,1] seed 2)
    to 1e-10, with the solution's residual recomputed against the oracle's
    operator_apply on the same slab."""
    from fehip.problems import MassProblem

    D, A = fehip.SCATTER_DETERMINISTIC, fehip.SCATTER_ATOMIC
    t1 = time.perf_counter()
    p5 = MassProblem(48, 8, 10, 3, scatter=D)
    fehip.synchronize()
    log("config 5 setup %.1f s" % (time.perf_counter() - t1))
    L5 = p5.l_size
    big = {"workload": "48^3 hexes, P=8, Q=10, num_comp=3, %d DoF" % L5}
    d_in = {np.float64: fehip.rng_uniform_device(1, L5, stream=stream)}
    u_host = d_in[np.float64].to_host()
    d_in[np.float32] = fehip.DeviceArray.from_host(u_host.astype(np.float32))
    d_out = {np.float64: fehip.DeviceArray(L5, np.float64), np.float32: fehip.DeviceArray(L5, np.float32)}
    steps = max(args.steps // 2, 10)
    # Each algorithm at its fastest packing on this operator (the autotuner's
    # candidates; for box columns also the component-split kernel).
    packing = {}
    for fp32, scatter in ((False, A), (True, D)):
        p5.op.set_fused_options(scatter, exact=0)
        for r in p5.op.autotune(fp32=fp32, reps=3)[1]:
            key = (fp32, r[0])
            if key not in packing or r[2] < packing[key][1]:
                packing[key] = (r[1], r[2])
    big["tuned_packings"] = {("fp32_" if k[0] else "fp64_") + fehip.FUSED_NAMES[k[1]]: v[0]
                             for k, v in packing.items()}
    for label, dt, scatter, alg in (("fp64_box_column", np.float64, D, fehip.FUSED_BOX_COLUMN),
                                    ("fp32_box_column", np.float32, D, fehip.FUSED_BOX_COLUMN),
                                    ("fp64_block_partial", np.float64, D, fehip.FUSED_BLOCK_PARTIAL),
                                    ("fp64_elem_tmap", np.float64, D, fehip.FUSED_ELEM_TMAP),
                                    ("fp64_block_atomic", np.float64, A, fehip.FUSED_BLOCK_ATOMIC),
                                    ("fp32_block_partial", np.float32, D, fehip.FUSED_BLOCK_PARTIAL)):
        p5.op.set_fused_options(scatter, exact=0)
        p5.op.set_fused_algorithm(alg, packing.get((dt == np.float32, alg), (0,))[0])
        m, k, a = time_applies(fehip, p5.op, d_in[dt], d_out[dt], steps, args.warmup, stream)
        sz = 4 if dt == np.float32 else 8
        ne5 = p5.op.fused_choice(dt == np.float32)[1]
        big[label] = {"ms_per_apply": m, "gdofs": L5 / m / 1e6, "kernel_ms": k, "other_stages_ms": a,
                      "elems": ne5,
                      "apply_frac_hbm": p5.apply_bytes(sz, sz, sz) / (m * 1e-3) / 1e9 / HBM_PEAK_GBS}
        if alg == fehip.FUSED_BOX_COLUMN:  # the kernel reads u and qdata and writes v (no offsets)
            kb = sz * (2 * L5 + p5.num_elem * p5.Q ** 3)
            big[label].update({"kernel": box_kernel_name(p5.P, p5.Q, p5.nc, dt == np.float32, ne5),
                               "kernel_bytes": kb, "kernel_frac_hbm": kb / (k * 1e-3) / 1e9 / HBM_PEAK_GBS})
        log("config5", label, json.dumps(big[label]))
    best = min((v["ms_per_apply"], k2) for k2, v in big.items() if k2.startswith("fp64"))[1]
    p5.op.set_fused_options(D, exact=0)
    best_alg = {"fp64_box_column": fehip.FUSED_BOX_COLUMN, "fp64_block_partial": fehip.FUSED_BLOCK_PARTIAL,
                "fp64_elem_tmap": fehip.FUSED_ELEM_TMAP}.get(best, fehip.FUSED_AUTO)
    p5.op.set_fused_algorithm(best_alg, packing.get((False, best_alg), (0,))[0])
    # The FP64 apply against the oracle on the slab.
    t1 = time.perf_counter()
    orc, take, l_s, rows = slab_oracle(p5, layers)
    p5.op.apply(d_in[np.float64], d_out[np.float64], stream)
    stream.synchronize()
    v_gpu = d_out[np.float64].to_host()
    v_cpu = orc.apply(take(u_host))
    big["apply_err_vs_cpu_slab"] = slab_rel_err(v_gpu, v_cpu, l_s, rows, 3, p5.num_nodes)
    big["cpu_slab"] = "oracle operator_apply on the first %d element layers (48x48x%d elements), rows below " \
                      "the slab's top node plane (%d nodes per component); %.1f s" % (
                          layers, layers, rows, time.perf_counter() - t1)
    # Solves.
    x_star = fehip.rng_uniform_device(2, L5, stream=stream)
    d_b, d_x = fehip.DeviceArray(L5), fehip.DeviceArray(L5)
    p5.op.apply(x_star, d_b, stream)
    stream.synchronize()
    xs, bh = x_star.to_host(), d_b.to_host()
    solves = {}
    for name, mode in (("fp64", fehip.SOLVE_FP64), ("mixed", fehip.SOLVE_MIXED)):
        p5.op.solve(d_b, d_x, mode=mode, rtol=1e-10)  # records the graphs
        lat = []
        for _ in range(3):
            t0 = time.perf_counter()
            st = p5.op.solve(d_b, d_x, mode=mode, rtol=1e-10)
            lat.append((time.perf_counter() - t0) * 1e3)
        x = d_x.to_host()
        Mx_cpu = orc.apply(take(x))
        b_s = take(bh)
        r = np.concatenate([b_s[c * l_s: c * l_s + rows] - Mx_cpu[c * l_s: c * l_s + rows] for c in range(3)])
        bn = np.concatenate([b_s[c * l_s: c * l_s + rows] for c in range(3)])
        solves[name] = {"p50_ms": float(np.median(lat)), "iterations": st["iterations"], "outer": st["outer"],
                        "rel_residual": st["rel_residual"],
                        "rel_residual_cpu_slab": float(np.linalg.norm(r) / np.linalg.norm(bn)),
                        "err_vs_xstar": float(np.max(np.abs(x - xs)) / np.max(np.abs(xs)))}
        log("config5 solve", name, json.dumps(solves[name]))
    solves["speedup_mixed_vs_fp64_p50"] = solves["fp64"]["p50_ms"] / solves["mixed"]["p50_ms"]
    big["solve"] = solves
    p5.op.set_fused_algorithm(fehip.FUSED_AUTO)
    return big


def solve_latency(fehip, reps, n=16, P=3, Q=4):
    """M x = b on the n^3 mesh (P, Q), b = M x*, x* ~ U[-1,1] seed 2, to a 1e-10
    relative residual: FP64 PCG, FP32-inner / FP64-refined PCG, and FP32 PCG with
    FP64 reliable updates. Latency percentiles over `reps` solves, iteration
    counts, and each solution's error against x* and against the FP64 solution."""
    from fehip.problems import MassProblem

    p = MassProblem(n, P, Q, 1)
    x_star = fehip.DeviceArray.from_host(fehip.rng_uniform(2, p.l_size))
    d_b, d_x = fehip.DeviceArray(p.l_size), fehip.DeviceArray(p.l_size)
    p.op.apply(x_star, d_b)
    xs = x_star.to_host()
    out = {"workload": "%d^3 hexes, P=%d, Q=%d, %d DoF, rtol 1e-10" % (n, P, Q, p.l_size)}
    x64 = None
    # fp64 / mixed run the automatic CG form (FP64: pipelined up to 2^17 unknowns,
    # plain above; FP32 inner solves plain; the plain form with its wide kernels
    # on vectors of 96 MiB and more); the last two lines force another form:
    # plain CG with the grid-stride kernels (FP64), pipelined (FP32 inner).
    for name, mode, cgv in (("fp64", fehip.SOLVE_FP64, fehip.CG_AUTO), ("mixed", fehip.SOLVE_MIXED, fehip.CG_AUTO),
                            ("mixed_reliable", fehip.SOLVE_MIXED_RELIABLE, fehip.CG_AUTO),
                            ("fp64_plain_cg", fehip.SOLVE_FP64, fehip.CG_PLAIN),
                            ("mixed_pipelined_cg", fehip.SOLVE_MIXED, fehip.CG_PIPELINED)):
        p.op.solve(d_b, d_x, mode=mode, rtol=1e-10, cg_variant=cgv)
        lat, st = [], None
        for _ in range(reps):
            t0 = time.perf_counter()
            st = p.op.solve(d_b, d_x, mode=mode, rtol=1e-10, cg_variant=cgv)
            lat.append((time.perf_counter() - t0) * 1e3)
        x = d_x.to_host()
        if x64 is None:
            x64 = x
        scale = float(np.max(np.abs(x64)))
        out[name] = {"p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
                     "iterations": st["iterations"], "outer": st["outer"], "rel_residual": st["rel_residual"],
                     "err_vs_xstar": float(np.max(np.abs(x - xs)) / np.max(np.abs(xs))),
                     "err_vs_fp64": float(np.max(np.abs(x - x64)) / scale)}
    out["speedup_mixed_vs_fp64_p50"] = out["fp64"]["p50_ms"] / out["mixed"]["p50_ms"]
    out["speedup_reliable_vs_fp64_p50"] = out["fp64"]["p50_ms"] / out["mixed_reliable"]["p50_ms"]
    out["speedup_auto_vs_plain_fp64_p50"] = out["fp64_plain_cg"]["p50_ms"] / out["fp64"]["p50_ms"]
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    # 200 applies of 0.25 ms after 100 untimed ones: on MI355X the first
    # milliseconds of load after an idle spell run up to 20% slower while the
    # clocks settle, which 5 warm-up applies do not cover.
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--n", type=int, default=64)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-solve", action="store_true")
    ap.add_argument("--no-big", action="store_true", help="skip the 48^3, P=8, Q=10, 3-component line")
    ap.add_argument("--no-serve", action="store_true", help="skip the serving workload")
    ap.add_argument("--serve-requests", type=int, default=2000)
    ap.add_argument("--serve-max-active", type=int, default=16)
    ap.add_argument("--serve-workers", type=int, default=8)
    ap.add_argument("--serve-unbatched", action="store_true",
                    help="one stream and captured graph per request instead of shape cohorts")
    ap.add_argument("--serve-rate", type=float, default=0.0,
                    help="Poisson arrival rate, requests/s (default: 0.8 x the measured service capacity)")
    ap.add_argument("--variants", default="", help="comma-separated subset of the apply variants")
    ap.add_argument("--box-layers", type=int, default=0,
                    help="most element layers per box-column segment (0: automatic)")
    ap.add_argument("--elems", type=int, default=0,
                    help="with --variants: the elems (packing) of the listed variants' algorithm (0: the default)")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="K",
                    help="also run K plain copies of one FP64 L-vector")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (--gpus must be 1)", file=sys.stderr)
        sys.exit(2)

    import fehip
    from fehip.problems import MassProblem

    info = fehip.device_info(0)
    log("device:", info.name.decode(), info.arch.decode(), "CUs", info.cu_count)
    log("tuning table rows loaded:", fehip.load_default_tuning())
    stream = fehip.Stream()
    n, P, Q = args.n, 5, 6
    t0 = time.perf_counter()
    p = MassProblem(n, P, Q, 1, scatter=fehip.SCATTER_DETERMINISTIC)
    if args.box_layers:
        p.op.set_box_layers(args.box_layers)
    fehip.synchronize()
    log("setup (mesh, build_mass on device): %.2f s" % (time.perf_counter() - t0))
    L = p.l_size

    # Every FP64 and FP32 fused-apply variant; the headline is the fastest FP64 one.
    bytes64 = p.apply_bytes()
    bytes32 = p.apply_bytes(4, 4, 4)
    E, P3 = p.num_elem, P ** 3
    variants = {}
    D, A = fehip.SCATTER_DETERMINISTIC, fehip.SCATTER_ATOMIC
    # The autotuner times every compiled algorithm and packing on this operator;
    # each variant below then runs at its algorithm's fastest packing.
    tuning = {}
    if not args.variants:
        for tag, scatter, fp32 in (("fp64", A, False), ("fp32", D, True)):
            p.op.set_fused_options(scatter, exact=0)
            best_t, rows = p.op.autotune(fp32=fp32, reps=10)
            tuning[tag] = {"best": {"alg": fehip.FUSED_NAMES[best_t[0]], "elems": best_t[1], "ms": best_t[2]},
                           "candidates": [{"alg": fehip.FUSED_NAMES[r[0]], "elems": r[1], "ms": r[2]} for r in rows]}
            log("autotune", tag, json.dumps(tuning[tag]["best"]))

    def tuned_elems(tag, alg):
        rows = [c for c in tuning.get(tag, {}).get("candidates", []) if c["alg"] == fehip.FUSED_NAMES[alg]]
        return min(rows, key=lambda c: c["ms"])["elems"] if rows else 0

    runs = (("fp64_box_column", np.float64, D, 0, fehip.FUSED_BOX_COLUMN),
            ("fp64_block_partial", np.float64, D, 0, fehip.FUSED_BLOCK_PARTIAL),
            ("fp64_elem_tmap", np.float64, D, 0, fehip.FUSED_ELEM_TMAP),
            ("fp64_block_partial_lds", np.float64, A, 0, fehip.FUSED_BLOCK_PARTIAL_LDS),
            ("fp64_block_atomic", np.float64, A, 0, fehip.FUSED_BLOCK_ATOMIC),
            ("fp64_elem_atomic", np.float64, A, 0, fehip.FUSED_ELEM_ATOMIC),
            ("fp64_exact_tmap", np.float64, D, 1, fehip.FUSED_ELEM_TMAP),
            ("fp32_box_column", np.float32, D, 0, fehip.FUSED_BOX_COLUMN),
            ("fp32_block_partial", np.float32, D, 0, fehip.FUSED_BLOCK_PARTIAL),
            ("fp32_elem_tmap", np.float32, D, 0, fehip.FUSED_ELEM_TMAP),
            ("fp32_block_atomic", np.float32, A, 0, fehip.FUSED_BLOCK_ATOMIC))
    VARIANT_ALG = {r[0]: r[4] for r in runs}
    DET_VARIANTS = {r[0] for r in runs if r[2] == D}
    wanted = set(filter(None, args.variants.split(",")))
    unknown = wanted - {r[0] for r in runs}
    if unknown:
        ap.error("unknown variants: %s" % ", ".join(sorted(unknown)))
    kernel_names = {}
    # Inputs made once, before any timing: u ~ U[-1,1] (splitmix64 seed 1)
    # generated on the device in FP64, its FP32 copy uploaded once; outputs
    # allocated once. Nothing runs on the host between the variants.
    d_in = {np.float64: fehip.rng_uniform_device(1, L, stream=stream), np.float32: None}
    d_in[np.float32] = fehip.DeviceArray.from_host(d_in[np.float64].to_host().astype(np.float32))
    d_out = {np.float64: fehip.DeviceArray(L, np.float64), np.float32: fehip.DeviceArray(L, np.float32)}
    for label, dt, scatter, exact, alg in runs:
        if wanted and label not in wanted:
            continue
        d_u, d_v = d_in[dt], d_out[dt]
        fp32 = dt == np.float32
        p.op.set_fused_options(scatter, exact=exact)
        if not exact:
            p.op.set_fused_algorithm(alg, args.elems or tuned_elems("fp32" if fp32 else "fp64", alg))
        ne = p.op.fused_choice(fp32)[1]
        tname = "float" if fp32 else "double"
        if alg in (fehip.FUSED_ELEM_ATOMIC, fehip.FUSED_ELEM_TMAP):
            kernel_names[label] = "fused_mass_%s_kernel<%d, %d, 1, %s, %d, %s>" % (
                "exact" if exact else "fast", P, Q, tname, ne, "true" if alg == fehip.FUSED_ELEM_TMAP else "false")
        elif alg == fehip.FUSED_BOX_COLUMN:
            kernel_names[label] = box_kernel_name(P, Q, 1, fp32, ne)
        else:
            kernel_names[label] = {
                fehip.FUSED_BLOCK_PARTIAL: "fused_mass_part_kernel<%d, %d, 1, %s, %d, false>",
                fehip.FUSED_BLOCK_PARTIAL_LDS: "fused_mass_part_kernel<%d, %d, 1, %s, %d, true>",
                fehip.FUSED_BLOCK_ATOMIC: "fused_mass_block_kernel<%d, %d, 1, %s, %d>"}[alg] % (P, Q, tname, ne)
        steps = args.steps if label.startswith("fp64") else max(args.steps // 2, 5)
        m, k, a = time_applies(fehip, p.op, d_u, d_v, steps, args.warmup, stream)
        s = 4 if fp32 else 8
        # The fused kernel's own algorithmic bytes: u, qdata and offsets read; v
        # written, except the element-output (tmap) kernels, which write E*P^3
        # values and leave the sums to the transpose-map pass, and the box-column
        # kernel, which reads no offsets (the box formula gives every node).
        tmap = alg == fehip.FUSED_ELEM_TMAP
        offs = 0 if alg == fehip.FUSED_BOX_COLUMN else 4 * E * P3
        kbytes = s * L + s * E * Q ** 3 + offs + (s * E * P3 if tmap else s * L)
        variants[label] = {"ms_per_apply": m, "gdofs": L / m / 1e6, "kernel_ms": k, "other_stages_ms": a,
                           "kernel_GBs": kbytes / (k * 1e-3) / 1e9, "kernel_frac_hbm": kbytes / (k * 1e-3) / 1e9 / HBM_PEAK_GBS,
                           "apply_frac_hbm": (bytes32 if s == 4 else bytes64) / (m * 1e-3) / 1e9 / HBM_PEAK_GBS,
                           "kernel_bytes": kbytes, "kernel": kernel_names[label]}
        log(label, json.dumps(variants[label]))
    p.op.set_fused_options(D, exact=0)
    if args.calibrate_copy:
        r = fehip.elem_restriction_create_strided(L, 1, 1, L)
        d_a = fehip.DeviceArray.from_host(fehip.rng_uniform(1, L))
        d_b = fehip.DeviceArray(L)
        for _ in range(args.calibrate_copy):
            fehip.elem_restriction_apply(r, fehip.NO_TRANSPOSE, d_a, d_b, stream)
        stream.synchronize()
        log("copy calibration: %d copies of %d bytes read + %d written" % (args.calibrate_copy, 8 * L, 8 * L))
    fp64 = [k2 for k2 in variants if k2.startswith("fp64") and "exact" not in k2]
    if not fp64:
        ap.error("--variants must include an FP64 non-exact variant")
    best = min((variants[k2] for k2 in fp64), key=lambda v: v["ms_per_apply"])
    best_name = [k2 for k2, v in variants.items() if v is best][0]
    ms = best["ms_per_apply"]

    big = None
    if not args.no_big:
        big = config5(fehip, args, stream)
        log("config5", json.dumps(big))

    solve = solve64 = None
    if not args.no_solve:
        solve = solve_latency(fehip, 20)  # config 3: the 16^3, P=3, Q=4 mesh of config 1
        log("solve", solve)
        solve64 = solve_latency(fehip, 5, 64, 5, 6)  # the headline mesh
        log("solve 64^3", solve64)
    serve = None
    if not args.no_serve:
        # Config 4: mass-solve requests (n ~ U{4..40}, P ~ U{2..6}, Q = P+1,
        # num_comp ~ U{1..3}, half FP64 and half mixed) with Poisson arrivals.
        from fehip.serving import make_requests, measure_capacity, run_workload

        t1 = time.perf_counter()
        opts = dict(max_active=args.serve_max_active, workers=args.serve_workers, hbm_fraction=0.5,
                    batched=not args.serve_unbatched)
        capacity = None
        rate = args.serve_rate
        if rate <= 0:
            capacity, _ = measure_capacity(300, **opts)
            rate = 0.8 * capacity
            log("serving capacity %.1f requests/s -> arrival rate %.1f/s" % (capacity, rate))
        serve, _ = run_workload(make_requests(args.serve_requests, seed=0), rate, **opts)
        serve["measured_capacity_req_per_s"] = capacity
        serve["server_options"] = opts
        # The same capacity measurement with one stream and captured graph per
        # request (no cohorts), for comparison.
        other = dict(opts, batched=not opts["batched"])
        cap2, s2 = measure_capacity(300, **other)
        serve["capacity_other_mode"] = {"batched": other["batched"], "capacity_req_per_s": cap2,
                                        "latency_ms": s2["latency_ms"], "setup_ms_p50": s2["setup_ms_p50"],
                                        "solve_ms_p50": s2["solve_ms_p50"]}
        serve["workload"] = ("%d mass-solve requests, Poisson arrivals at %.1f/s (0.8 x measured capacity), "
                             "n ~ U{4..40}, P ~ U{2..6}, Q = P+1, num_comp ~ U{1..3}, FP64 or mixed (1:1), "
                             "rtol 1e-10, request r seeded 1000+r; %s" % (
                                 args.serve_requests, rate,
                                 "shape cohorts: concurrent requests of one (P, Q, num_comp, precision) share each "
                                 "persistent element/step launch of one captured graph" if opts["batched"] else
                                 "one stream and captured graph per request"))
        log("serve (%.1f s)" % (time.perf_counter() - t1), json.dumps(serve))
    cpu = None
    if not args.no_cpu:
        # The headline variant's output on the same u, for its error against the oracle.
        p.op.set_fused_options(D if best_name in DET_VARIANTS else A, exact=0)
        p.op.set_fused_algorithm(VARIANT_ALG[best_name], tuned_elems("fp64", VARIANT_ALG[best_name]))
        p.op.apply(d_in[np.float64], d_out[np.float64], stream)
        stream.synchronize()
        cpu = cpu_baseline(p, d_out[np.float64].to_host())
        log("cpu", cpu)

    k_ms = best["kernel_ms"]
    achieved = best["kernel_bytes"] / (k_ms * 1e-3) / 1e9
    copy = copy_ceiling(fehip, stream)
    log("streaming copy ceiling:", json.dumps(copy))
    line = {
        "metric": "mass-apply GDoF/s & %HBM roofline, FP64 vs mixed; p50/p99 solve latency",
        "value": L / ms / 1e6,
        "unit": "GDoF/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (jittered structured hex mesh, u ~ U[-1,1] splitmix64 seed 1)",
        "config": {"workload": "fused FP64 mass operator_apply, 64^3 hexes, P=5 GLL, Q=6 Gauss, num_comp=1, "
                               "%d DoF, %s scatter" % (L, best_name.split("_", 1)[1].replace("_", " ")),
                   "num_elem": p.num_elem, "P": P, "Q": Q, "num_comp": 1, "l_size": L},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic_from_profile(best["kernel"]),
                     "kernel": best["kernel"],
                     "kernel_ms": k_ms, "algorithmic_bytes_per_launch": best["kernel_bytes"],
                     "apply_algorithmic_bytes": bytes64, "apply_frac": best["apply_frac_hbm"],
                     "other_stages_ms": best["other_stages_ms"],
                     "copy_GBs": copy["GBs"], "frac_of_copy": achieved / copy["GBs"]},
        "cpu_baseline": cpu,
        "variants": variants,
        "tuning": tuning,
        "config5": big,
        "solve_c3": solve,
        "solve_64cube": solve64,
        "serving_c4": serve,
        "device": {"name": info.name.decode(), "arch": info.arch.decode(), "cu": info.cu_count},
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
