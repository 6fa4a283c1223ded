"""Box-column algorithm (FUSED_BOX_COLUMN) of the fused mass apply vs the oracle.

The algorithm exists only for structured boxes (lexicographic node numbering),
which the library recognises from the offsets themselves. On rectangular boxes
with every compiled column cross-section and several layer counts per
workgroup (so that segment ends, column faces, the shared-node pass and
domain-boundary faces are all exercised), FP64 agrees with the oracle's
operator_apply within 1e-13 relative and FP32 within 2e-6; outputs start as
NaN (an unwritten node shows), and two applies agree bit for bit. Offsets that
are not such a box (permuted numbering) never get the algorithm.
"""
import numpy as np
import pytest

import fehip
from fehip.meshes import MeshProblem, box_mesh, permute
from helpers import rel_err
from oracle_lib import oracle

pytestmark = pytest.mark.gpu

FP64_RTOL = 1e-13
BOX = fehip.FUSED_BOX_COLUMN


def rect_mesh(nx, ny, nz, P, jitter=0.1, seed=3):
    """nx*ny*nz hexes of the unit cube with interior vertices jittered (numpy RNG)."""
    g = [np.linspace(0.0, 1.0, n + 1) for n in (nx, ny, nz)]
    z, y, x = np.meshgrid(g[2], g[1], g[0], indexing="ij")
    coords = np.stack([x.ravel(), y.ravel(), z.ravel()])
    rng = np.random.default_rng(seed)
    iz, iy, ix = np.meshgrid(np.arange(nz + 1), np.arange(ny + 1), np.arange(nx + 1), indexing="ij")
    inner = ((ix > 0) & (ix < nx) & (iy > 0) & (iy < ny) & (iz > 0) & (iz < nz)).ravel()
    h = np.array([1.0 / nx, 1.0 / ny, 1.0 / nz])[:, None]
    coords[:, inner] += jitter * h * rng.uniform(-1, 1, (3, inner.sum()))
    Nn = [n * (P - 1) + 1 for n in (nx, ny, nz)]
    return {"dim": 3, "num_elem": nx * ny * nz, "P": P, "num_nodes": Nn[0] * Nn[1] * Nn[2],
            "num_vertices": (nx + 1) * (ny + 1) * (nz + 1), "offsets": fehip.box_offsets(nx, ny, nz, P, 1),
            "vertex_offsets": fehip.box_offsets(nx, ny, nz, 2, 1), "coords": coords.ravel()}


# (box, P, Q, num_comp, elems per column layer, most layers per workgroup to
# try): every cross-section at P=5, and one compiled variant of each other
# shape (P = 2, 3, 4, 6, 8; 1 to 3 components). Limits that do not divide nz
# give segments of unequal length (8 layers at most 3: 2, 3, 3).
CASES = [((8, 8, 8), 5, 6, 1, 16, (0, 1, 2, 3, 8)),
         ((8, 8, 6), 5, 6, 1, 8, (0, 3, 4)),
         ((4, 6, 6), 5, 6, 1, 4, (0, 2, 4, 6)),
         ((4, 6, 4), 5, 6, 3, 8, (0, 2, 3)),
         ((4, 4, 5), 2, 3, 3, 8, (0, 3)),
         ((8, 4, 4), 2, 3, 1, 16, (0, 2)),
         ((4, 4, 7), 3, 4, 2, 8, (0, 1, 3)),
         ((4, 4, 4), 4, 5, 3, 4, (0, 2, 3)),
         ((4, 4, 3), 6, 7, 1, 8, (0, 2)),
         ((2, 2, 5), 8, 10, 3, 4, (0, 2)),
         ((4, 4, 3), 8, 10, 3, 8, (0, 2)),
         ((4, 2, 2), 8, 10, 1, 4, (0,)),
         ((16, 8, 4), 5, 6, 1, 32, (0, 2, 4)),
         ((16, 4, 5), 5, 6, 1, 18, (0, 2, 3)),
         ((32, 4, 3), 5, 6, 1, 34, (0, 2)),
         ((16, 4, 3), 8, 10, 3, 18, (0, 2))]
# Shapes whose cross-section is compiled only as the component-split kernel.
SPLIT_ONLY = {(8, 10, 3, 8)}
# Cross-sections compiled only as wavefront kernels (codes: 8 x 4 is 32, 8 x 2
# is 18, 16 x 2 is 34).
WAVE_ONLY = {(5, 6, 1, 32): (632,), (5, 6, 1, 18): (618,), (5, 6, 1, 34): (634,), (8, 10, 3, 18): (518,)}
# The kernels with whole wavefronts per element compiled (elems + 400: one
# element per wavefront, + 600: two; + 100 for the component split of a
# multi-component field).
WAVE = {(5, 6, 1, 16): (616,), (5, 6, 1, 8): (608,), (5, 6, 3, 8): (708,), (2, 3, 3, 8): (708,),
        (2, 3, 1, 16): (616,), (3, 4, 2, 8): (708,), (4, 5, 3, 4): (704,), (6, 7, 1, 8): (408,),
        (8, 10, 3, 8): (508,), (8, 10, 3, 4): (504,), (8, 10, 1, 4): (404,)}


@pytest.mark.parametrize("n3,P,Q,nc,elems,layers", CASES)
def test_box_column_matches_oracle(gpu, n3, P, Q, nc, elems, layers):
    print("box column vs oracle", n3, P, Q, nc, elems, layers, flush=True)
    mesh = rect_mesh(*n3, P)
    p = MeshProblem(mesh, Q, nc, scatter=fehip.SCATTER_DETERMINISTIC)
    N = mesh["num_nodes"]
    assert fehip.box_detect(mesh["num_elem"], P, mesh["offsets"], nc, N, nc * N) == n3
    want = oracle().mass(3, mesh["num_elem"], P, Q, nc, nc * N, N, mesh["offsets"], p.qdata()).apply(
        fehip.rng_uniform(5, nc * N))
    u = fehip.rng_uniform(5, nc * N)
    d_u, d_u32 = fehip.DeviceArray.from_host(u), fehip.DeviceArray.from_host(u.astype(np.float32))
    # elems + 100: the kernel that loads qdata within its own layer; + 200: FP64
    # through the LDS ring (FP32 has no such variant: it always takes the ring);
    # + 300 (several components): one component per workgroup; + 400 / + 600:
    # one / two elements per wavefront (+ 100: with the component split).
    variants = (elems, elems + 100, elems + 200) + ((elems + 300,) if nc > 1 else ())
    if (P, Q, nc, elems) in SPLIT_ONLY:
        variants = (elems + 300,)
    if (P, Q, nc, elems) in WAVE:
        variants += WAVE[(P, Q, nc, elems)]
    if (P, Q, nc, elems) in WAVE_ONLY:
        variants = WAVE_ONLY[(P, Q, nc, elems)]
    for lz, el in [(lz, el) for lz in layers for el in variants]:
        has32 = el // 100 != 2
        fehip.box_column_check(*n3, P, elems, lz)  # the index arithmetic, on the host
        p.op.set_box_layers(lz)
        p.op.set_fused_algorithm(BOX, el)
        assert p.op.fused_choice() == (BOX, el)
        if has32:
            assert p.op.fused_choice(fp32=True) == (BOX, el)
        d_v = fehip.DeviceArray.from_host(np.full(nc * N, np.nan))
        p.op.apply(d_u, d_v)
        got = d_v.to_host()
        assert rel_err(got, want) < FP64_RTOL, (lz, el)
        d_v = fehip.DeviceArray.from_host(np.full(nc * N, np.nan))
        p.op.apply(d_u, d_v)
        assert d_v.to_host().tobytes() == got.tobytes(), (lz, el)
        if not has32:
            continue
        d_v32 = fehip.DeviceArray.from_host(np.full(nc * N, np.nan, np.float32))
        p.op.apply(d_u32, d_v32)
        assert rel_err(d_v32.to_host().astype(np.float64), want) < 2e-6, (lz, el)


def test_box_column_is_the_default_on_large_boxes(gpu):
    """Without a tuning row, a deterministic box operator of at least 1 Mi
    element nodes runs the box columns, a smaller one does not (too few column
    segments to fill the GPU); a layer limit that does not divide nz keeps
    them (segments of unequal length)."""
    if any(r[:5] == (5, 6, 1, 0, 1) for r in fehip.tune_table()):  # any (P, Q, nc, FP64, det) row
        pytest.skip("a tuning row for this shape is loaded")
    small = MeshProblem(rect_mesh(8, 8, 8, 5), 6, 1, scatter=fehip.SCATTER_DETERMINISTIC)
    assert small.op.fused_choice()[0] != BOX
    small.op.set_fused_algorithm(BOX)
    assert small.op.fused_choice()[0] == BOX
    del small
    mesh = rect_mesh(24, 24, 24, 5)  # 13,824 elements, 1.7M element nodes
    p = MeshProblem(mesh, 6, 1, scatter=fehip.SCATTER_DETERMINISTIC)
    assert p.op.fused_choice()[0] == BOX
    p.op.set_box_layers(5)  # 5 segments of 4 or 5 layers
    assert p.op.fused_choice()[0] == BOX


def test_box_column_refused_off_box(gpu):
    """Permuted node numbering: not a box, so never the box columns."""
    mesh = permute(box_mesh(4, 5), seed=2)
    assert fehip.box_detect(mesh["num_elem"], 5, mesh["offsets"]) is None
    p = MeshProblem(mesh, 6, 1, scatter=fehip.SCATTER_DETERMINISTIC)
    assert p.op.fused_choice()[0] != BOX
    p.op.set_fused_algorithm(BOX)
    with pytest.raises(ValueError):
        p.op.fused_choice()
