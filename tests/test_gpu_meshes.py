"""The fused mass operator on meshes that are not the lexicographic box: a
permuted L-shaped domain (any element order and node numbering, nodes on the
re-entrant edge touching 6 elements), and the committed FEMHMSH1 fixture.

* exact arithmetic + deterministic scatter equals the oracle bit for bit;
* every scatter algorithm (element/block, atomic/partial) within 1e-13, FP32
  within 2e-6 (the block plans cluster elements from the offsets alone);
* 1^T M 1 = 3/4; FP64 and mixed solves reach 1e-10.
"""
import os

import numpy as np
import pytest

import fehip
from fehip.meshes import MeshProblem, l_shape, permute
from helpers import rel_err
from oracle_lib import oracle

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
ALGS = [fehip.FUSED_ELEM_ATOMIC, fehip.FUSED_ELEM_TMAP, fehip.FUSED_BLOCK_ATOMIC, fehip.FUSED_BLOCK_PARTIAL,
        fehip.FUSED_BLOCK_PARTIAL_LDS]
DET = {fehip.FUSED_ELEM_TMAP, fehip.FUSED_BLOCK_PARTIAL}


@pytest.mark.parametrize("n,P,Q", [(4, 3, 4), (6, 5, 6), (4, 2, 3)])
def test_permuted_l_shape_matches_oracle(gpu, n, P, Q):
    print("l-shape", n, P, Q, flush=True)
    m = permute(l_shape(n, P), seed=5)
    p = MeshProblem(m, Q)
    o = oracle()
    E, N, nv = m["num_elem"], m["num_nodes"], m["num_vertices"]
    qd = o.build_qdata(3, E, 2, Q, 3 * nv, nv, m["vertex_offsets"], m["coords"])
    assert p.qdata().tobytes() == qd.tobytes()
    assert abs(qd.sum() - 0.75) < 1e-12
    u = fehip.rng_uniform(1, N)
    want = o.mass(3, E, P, Q, 1, N, N, m["offsets"], qd).apply(u)
    d_u = fehip.DeviceArray.from_host(u)
    d_v = fehip.DeviceArray(N)
    p.op.set_fused_options(fehip.SCATTER_DETERMINISTIC, exact=1)
    p.op.apply(d_u, d_v)
    assert d_v.to_host().tobytes() == want.tobytes()
    d_u32 = fehip.DeviceArray.from_host(u.astype(np.float32))
    for alg in ALGS:
        p.op.set_fused_options(fehip.SCATTER_DETERMINISTIC if alg in DET else fehip.SCATTER_ATOMIC, exact=0)
        p.op.set_fused_algorithm(alg)
        d_v = fehip.DeviceArray.from_host(np.full(N, np.nan))
        p.op.apply(d_u, d_v)
        assert rel_err(d_v.to_host(), want) < 1e-13, alg
        d_v32 = fehip.DeviceArray(N, np.float32)
        p.op.apply(d_u32, d_v32)
        assert rel_err(d_v32.to_host().astype(np.float64), want) < 2e-6, alg


def test_committed_fixture_on_device(gpu):
    path = os.path.join(HERE, "golden", "lshape_n4_p3.femh")
    ref = np.load(os.path.join(HERE, "golden", "lshape_n4_p3.npz"))
    p = MeshProblem(path, 4)
    assert p.qdata().tobytes() == ref["qdata"].tobytes()
    d_u, d_v = fehip.DeviceArray.from_host(ref["u"]), fehip.DeviceArray(p.l_size)
    p.op.set_fused_options(fehip.SCATTER_DETERMINISTIC, exact=1)
    p.op.apply(d_u, d_v)
    assert d_v.to_host().tobytes() == ref["v"].tobytes()


def test_solves_on_irregular_mesh(gpu):
    m = permute(l_shape(8, 4), seed=9)
    p = MeshProblem(m, 5)
    xs = fehip.DeviceArray.from_host(fehip.rng_uniform(2, p.l_size))
    b, x = fehip.DeviceArray(p.l_size), fehip.DeviceArray(p.l_size)
    p.op.apply(xs, b)
    for mode in (fehip.SOLVE_FP64, fehip.SOLVE_MIXED):
        st = p.op.solve(b, x, mode=mode, rtol=1e-10)
        assert st["rel_residual"] <= 1e-10, (mode, st)
        assert rel_err(x.to_host(), xs.to_host()) < 1e-7
