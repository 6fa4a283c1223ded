#!/usr/bin/env python3
"""Autotunes every compiled fused-mass shape on this GPU and writes the table.

For each (P, Q, num_comp) with a compiled fused kernel and each of six mesh
sizes (jittered structured hex meshes of about 180M, 34M, 24M, 6M, 1.5M and
0.375M element nodes, at most 96^3 hexes: the largest are far larger than the
caches, 180M is 48^3 at P=8 with 3 components and 34M is 64^3 at P=5, the
smallest launch-latency-bound), the autotuner times every compiled algorithm
and packing: FP64 with the atomic scatter allowed (every
algorithm; the fastest deterministic one is recorded as the FP64
deterministic choice) and FP32 with the deterministic scatter (the mixed
solve's inner apply). Each choice is recorded twice: for structured boxes
(box = 1, the box-column algorithm included where it is compiled) and for any
node numbering (box = 0, the best algorithm other than the box columns, as
timed on the box). An operator follows the rows measured on the size
nearest its own. The table is written as text (default: mi355x_gfx950.txt
beside this script), which fehip.load_default_tuning() and
fem_hip_tune_table_load() read.

Usage: python fe-demo-cpp_amd/tuning/tune_all.py [--out PATH] [--reps N] [--shapes P,Q,NC;...]
(--shapes: re-measure only those shapes; the other rows are kept from the
table at --out.)
"""
import argparse
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))

import fehip  # noqa: E402
from fehip.problems import MassProblem  # noqa: E402

SHAPES = [(2, 3, (1, 2, 3)), (3, 4, (1, 2, 3)), (4, 5, (1, 2, 3)), (5, 6, (1, 2, 3)), (6, 7, (1, 2, 3)),
          (8, 10, (1, 3))]
DET = (fehip.FUSED_ELEM_TMAP, fehip.FUSED_BLOCK_PARTIAL, fehip.FUSED_BOX_COLUMN)


TARGETS = (375_000, 1_500_000, 6_000_000, 24_000_000, 34_000_000, 180_000_000)  # element nodes; the largest last (it sets the alg rows)


def mesh_size(P, nc, target, n_max=96):
    """Hexes per side: about `target` element nodes (all components), a
    multiple of 4 (so the box-column cross-sections tile the box)."""
    n = 4
    while n < n_max and (n + 4) ** 3 * P ** 3 * nc <= target:
        n += 4
    return n


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=os.path.join(HERE, "mi355x_gfx950.txt"))
    ap.add_argument("--reps", type=int, default=8)
    ap.add_argument("--shapes", default="", help="P,Q,NC;... to re-measure (others kept from --out)")
    args = ap.parse_args()
    info = fehip.device_info(0)
    fehip.tune_table_clear()
    only = {tuple(int(x) for x in t.split(",")) for t in args.shapes.split(";") if t}
    if only:
        fehip.tune_table_load(args.out)
    for P, Q, ncs in SHAPES:
        for nc, n in sorted({(nc, mesh_size(P, nc, t)) for nc in ncs for t in TARGETS}):
            if only and (P, Q, nc) not in only:
                continue
            t0 = time.perf_counter()
            p = MassProblem(n, P, Q, nc, scatter=fehip.SCATTER_DETERMINISTIC)
            p.op.set_fused_options(fehip.SCATTER_ATOMIC, exact=0)
            (alg, ne, ms), rows = p.op.autotune(fp32=False, reps=args.reps)
            p.op.set_fused_options(fehip.SCATTER_DETERMINISTIC, exact=0)
            _, rows32 = p.op.autotune(fp32=True, reps=args.reps)
            picks = {}
            for fp32, det, cands in ((0, 0, rows), (0, 1, rows), (1, 1, rows32)):
                if det:
                    cands = [r for r in cands if r[0] in DET]
                for box in (1, 0):
                    pool = cands if box else [r for r in cands if r[0] != fehip.FUSED_BOX_COLUMN]
                    if not pool or (box and not any(r[0] == fehip.FUSED_BOX_COLUMN for r in cands)):
                        continue
                    a, e, t = min(pool, key=lambda r: r[2])
                    fehip.tune_table_set(P, Q, nc, fp32, det, a, e, t, n ** 3, box=box)
                    picks[(fp32, det, box)] = "%s/%d %.4f ms" % (fehip.FUSED_NAMES[a], e, t)
            print("P=%d Q=%d nc=%d n=%d  %s" % (P, Q, nc, n, "  ".join(
                "%s%s%s: %s" % ("fp32" if k[0] else "fp64", " det" if k[1] else "", " box" if k[2] else "", v)
                for k, v in sorted(picks.items()))), flush=True)
            del p
            print("  (%.1f s)" % (time.perf_counter() - t0), flush=True)
    fehip.tune_table_save(args.out)
    with open(args.out) as f:
        body = f.read()
    with open(args.out, "w") as f:
        f.write("# fused-mass tuning table, %s (%s, %d CUs); written by tune_all.py\n" % (
            info.arch.decode(), info.name.decode() or "MI355X", info.cu_count))
        f.write(body)
    print("wrote", args.out)


if __name__ == "__main__":
    main()
