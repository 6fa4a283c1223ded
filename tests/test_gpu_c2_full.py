"""The headline workload at full size against the oracle: 64^3 jittered hexes,
P=5 GLL nodes, Q=6 Gauss points, one component (16,974,593 DoF).

The oracle (the CPU restatement of the reference's operator_apply,
src/operator.cpp:148-182) builds qdata and applies the mass operator once on the
same seeded inputs (about 20 s on one core); then
  * the device build_mass is bit-equal to the oracle's qdata;
  * exact mode with the deterministic transpose-map scatter is bit-equal to the
    oracle's apply (the CPU scatter's sum order, src/elem-restriction.cpp:73-81);
  * every other FP64 algorithm (box columns, element + transpose map, block
    partials, atomics) is within 1e-13 relative, and the deterministic ones repeat
    bit for bit;
  * the FP32 applies are within 2e-6 relative of the FP64 oracle.
"""
import numpy as np
import pytest

import fehip
from fehip.problems import MassProblem
from helpers import rel_err
from oracle_lib import oracle

pytestmark = pytest.mark.gpu

N, P, Q = 64, 5, 6


@pytest.fixture(scope="module")
def c2():
    print("64^3 P=5 Q=6: device setup", flush=True)
    p = MassProblem(N, P, Q, 1, scatter=fehip.SCATTER_DETERMINISTIC)
    nv = (N + 1) ** 3
    print("64^3: oracle build_qdata", flush=True)
    q_orc = oracle().build_qdata(3, p.num_elem, 2, Q, 3 * nv, nv, p.mesh_offsets, p.coords)
    u = fehip.rng_uniform(1, p.l_size)
    print("64^3: oracle operator_apply", flush=True)
    want = oracle().mass(3, p.num_elem, P, Q, 1, p.l_size, p.num_nodes, p.offsets, q_orc).apply(u)
    print("64^3: oracle done", flush=True)
    return p, q_orc, u, want


def test_c2_qdata_bit_equal(gpu, c2):
    p, q_orc, _, _ = c2
    assert p.qdata().tobytes() == q_orc.tobytes()


def test_c2_exact_transpose_map_bit_equal(gpu, c2):
    p, _, u, want = c2
    d_u, d_v = fehip.DeviceArray.from_host(u), fehip.DeviceArray.from_host(np.full(p.l_size, np.nan))
    p.op.set_fused_options(fehip.SCATTER_DETERMINISTIC, exact=1)
    try:
        p.op.apply(d_u, d_v)
        got = d_v.to_host()
        assert got.tobytes() == want.tobytes(), rel_err(got, want)
    finally:
        p.op.set_fused_options(fehip.SCATTER_DETERMINISTIC, exact=0)


@pytest.mark.parametrize("alg", [fehip.FUSED_BOX_COLUMN, fehip.FUSED_ELEM_TMAP, fehip.FUSED_BLOCK_PARTIAL,
                                 fehip.FUSED_ELEM_ATOMIC, fehip.FUSED_BLOCK_ATOMIC])
def test_c2_fp64_algorithms(gpu, c2, alg):
    p, _, u, want = c2
    det = alg in (fehip.FUSED_BOX_COLUMN, fehip.FUSED_ELEM_TMAP, fehip.FUSED_BLOCK_PARTIAL)
    p.op.set_fused_options(fehip.SCATTER_DETERMINISTIC if det else fehip.SCATTER_ATOMIC, exact=0)
    p.op.set_fused_algorithm(alg)
    assert p.op.fused_choice()[0] == alg
    d_u, d_v = fehip.DeviceArray.from_host(u), fehip.DeviceArray.from_host(np.full(p.l_size, np.nan))
    p.op.apply(d_u, d_v)
    got = d_v.to_host()
    err = rel_err(got, want)
    print("alg", alg, "rel err", err, flush=True)
    assert err < 1e-13, (alg, err)
    if det:
        p.op.apply(d_u, d_v)
        assert d_v.to_host().tobytes() == got.tobytes(), alg
    p.op.set_fused_algorithm(fehip.FUSED_AUTO)


@pytest.mark.parametrize("alg", [fehip.FUSED_BOX_COLUMN, fehip.FUSED_BLOCK_PARTIAL])
def test_c2_fp32_algorithms(gpu, c2, alg):
    p, _, u, want = c2
    p.op.set_fused_options(fehip.SCATTER_DETERMINISTIC, exact=0)
    p.op.set_fused_algorithm(alg)
    d_u = fehip.DeviceArray.from_host(u.astype(np.float32))
    d_v = fehip.DeviceArray.from_host(np.full(p.l_size, np.nan, np.float32))
    p.op.apply(d_u, d_v)
    err = rel_err(d_v.to_host().astype(np.float64), want)
    print("fp32 alg", alg, "rel err", err, flush=True)
    assert err < 2e-6, (alg, err)
    p.op.set_fused_algorithm(fehip.FUSED_AUTO)
