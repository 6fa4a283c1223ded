"""Device element restriction vs the CPU oracle: the cases of the reference's
tests/cuda/test-cuda-elem-restriction.cu, with the deterministic scatter held to
bit equality (it starts from the current L value and sums in ascending t, the
CPU order of src/elem-restriction.cpp:73-81)."""
import numpy as np
import pytest

import fehip
from helpers import box_restriction_args, wiggly
from oracle_lib import oracle

pytestmark = pytest.mark.gpu

MODES = [fehip.SCATTER_ATOMIC, fehip.SCATTER_DETERMINISTIC]


def check_against_cpu(ne, es, nc, cs, l_size, offsets, mode):
    """tests/cuda/test-cuda-elem-restriction.cu:60-79, tightened: gather and the
    deterministic scatter are bit-exact; the atomic scatter is within 1e-13."""
    orc = oracle()
    r = fehip.elem_restriction_create(ne, es, nc, cs, l_size, offsets, mode)
    l_in = wiggly(l_size, 0.1)
    e_cpu = orc.restriction_apply(ne, es, nc, cs, l_size, offsets, 0, l_in)
    d_l, d_e = fehip.DeviceArray.from_host(l_in), fehip.DeviceArray(nc * es * ne)
    fehip.elem_restriction_apply(r, fehip.NO_TRANSPOSE, d_l, d_e)
    assert d_e.to_host().tobytes() == e_cpu.tobytes()

    e_in = wiggly(nc * es * ne, 0.7)
    l_cpu = orc.restriction_apply(ne, es, nc, cs, l_size, offsets, 1, e_in, out=np.ones(l_size))
    d_ein, d_lout = fehip.DeviceArray.from_host(e_in), fehip.DeviceArray.from_host(np.ones(l_size))
    fehip.elem_restriction_apply(r, fehip.TRANSPOSE, d_ein, d_lout)
    got = d_lout.to_host()
    if mode == fehip.SCATTER_DETERMINISTIC:
        assert got.tobytes() == l_cpu.tobytes()
    else:
        np.testing.assert_allclose(got, l_cpu, rtol=0, atol=1e-13)


@pytest.mark.parametrize("mode", MODES)
def test_small_cases(gpu, mode):
    print("restriction small cases, mode", mode, flush=True)
    check_against_cpu(2, 2, 1, 1, 3, np.array([0, 1, 1, 2]), mode)
    check_against_cpu(2, 2, 1, 1, 6, np.array([2, 4, 5, 2]), mode)  # unused L-nodes keep their value
    check_against_cpu(2, 4, 2, 6, 12, np.array([0, 1, 3, 4, 1, 2, 4, 5]), mode)


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("interleaved", [False, True])
@pytest.mark.parametrize("shape", [(4, 3, 2, 2, 1), (3, 3, 3, 3, 3), (20, 20, 20, 3, 1)])
def test_3d_meshes(gpu, mode, interleaved, shape):
    print("restriction 3D", shape, "interleaved", interleaved, "mode", mode, flush=True)
    check_against_cpu(*box_restriction_args(*shape, interleaved), mode)


def test_deterministic_bitwise_reproducible(gpu):
    """tests/cuda/test-cuda-elem-restriction.cu:211-233, plus equality with the CPU."""
    print("deterministic reproducibility 16^3 P=3 nc=3", flush=True)
    ne, es, nc, cs, l_size, off = box_restriction_args(16, 16, 16, 3, 3, True)
    r = fehip.elem_restriction_create(ne, es, nc, cs, l_size, off, fehip.SCATTER_DETERMINISTIC)
    e = wiggly(nc * es * ne, 0.3)
    d_e = fehip.DeviceArray.from_host(e)
    first = None
    for _ in range(5):
        d_l = fehip.DeviceArray(l_size)
        d_l.zero()
        fehip.elem_restriction_apply(r, fehip.TRANSPOSE, d_e, d_l)
        got = d_l.to_host()
        if first is None:
            first = got
        else:
            assert got.tobytes() == first.tobytes()
    cpu = oracle().restriction_apply(ne, es, nc, cs, l_size, off, 1, e, out=np.zeros(l_size))
    assert first.tobytes() == cpu.tobytes()


def transpose_map_cases():
    """(name, ne, es, nc, cs, l_size, offsets): a small interleaved box, a
    permuted L-shape (every node used, irregular valence), the same with every
    other L-node unused, a 0.9M-entry box with blocked components, and a hub
    node that 20,000 entries map to (one very long row)."""
    from fehip.meshes import l_shape, permute
    yield ("box interleaved",) + tuple(box_restriction_args(2, 2, 1, 2, 3, True))
    m = permute(l_shape(6, 3), seed=4)
    E, N = m["num_elem"], m["num_nodes"]
    yield "permuted l-shape", E, 27, 1, N, N, np.asarray(m["offsets"], np.int32)
    yield "unused nodes", E, 27, 2, 1, 2 * N + 1, (2 * np.asarray(m["offsets"])).astype(np.int32)
    n, P = 24, 4
    N = (n * (P - 1) + 1) ** 3
    yield "box 24^3 P=4", n ** 3, P ** 3, 3, N, 3 * N, fehip.box_offsets(n, n, n, P, 1)
    rng = np.random.default_rng(11)
    ne, es, l_size = 20000, 8, 5000
    off = rng.integers(1, l_size, ne * es).astype(np.int32)
    off[::es] = 0
    yield "hub node", ne, es, 1, l_size, l_size, off


@pytest.mark.parametrize("case", list(range(5)))
def test_transpose_map_on_device_matches_oracle(gpu, case):
    """The deterministic restriction builds its map on the device (atomic
    counts, scans, atomic fill, per-row sort); it equals the CPU counting sort
    entry for entry."""
    name, ne, es, nc, cs, l_size, off = list(transpose_map_cases())[case]
    print("device transpose map:", name, flush=True)
    r = fehip.elem_restriction_create(ne, es, nc, cs, l_size, off, fehip.SCATTER_DETERMINISTIC)
    lv, to, ti = r.transpose_map()
    lv2, to2, ti2 = oracle().transpose_map(ne, es, l_size, off)
    assert len(lv) == len(lv2) and (lv == lv2).all() and (to == to2).all() and (ti == ti2).all()


def test_strided(gpu):
    print("strided restrictions", flush=True)
    orc = oracle()
    # backend strides: gather is a copy
    ne, es, nc = 5, 27, 7
    r = fehip.elem_restriction_create_strided(ne, es, nc, ne * es * nc)
    l_in = wiggly(ne * es * nc, 0.1)
    d_l, d_e = fehip.DeviceArray.from_host(l_in), fehip.DeviceArray(ne * es * nc)
    fehip.elem_restriction_apply(r, fehip.NO_TRANSPOSE, d_l, d_e)
    assert d_e.to_host().tobytes() == l_in.tobytes()
    # user strides [e][i][k]
    ne, es, nc = 4, 8, 3
    strides = [nc, 1, es * nc]
    r = fehip.elem_restriction_create_strided(ne, es, nc, ne * es * nc, strides)
    l_in = wiggly(ne * es * nc, 0.2)
    d_l, d_e = fehip.DeviceArray.from_host(l_in), fehip.DeviceArray(ne * es * nc)
    fehip.elem_restriction_apply(r, fehip.NO_TRANSPOSE, d_l, d_e)
    e_cpu = orc.restriction_apply(ne, es, nc, 0, ne * es * nc, None, 0, l_in, strides=strides)
    assert d_e.to_host().tobytes() == e_cpu.tobytes()
    e_in = wiggly(ne * es * nc, 0.9)
    d_ein, d_lo = fehip.DeviceArray.from_host(e_in), fehip.DeviceArray.from_host(np.ones(ne * es * nc))
    fehip.elem_restriction_apply(r, fehip.TRANSPOSE, d_ein, d_lo)
    l_cpu = orc.restriction_apply(ne, es, nc, 0, ne * es * nc, None, 1, e_in, out=np.ones(ne * es * nc),
                                  strides=strides)
    assert d_lo.to_host().tobytes() == l_cpu.tobytes()


def test_apply_checks_sizes(gpu):
    """tests/cuda/test-cuda-elem-restriction.cu:158-171"""
    r = fehip.elem_restriction_create(2, 2, 1, 1, 3, np.array([0, 1, 1, 2]))
    out = fehip.DeviceArray(4)
    with pytest.raises(fehip.InvalidArgument):
        fehip.elem_restriction_apply(r, fehip.NO_TRANSPOSE, fehip.DeviceArray(4), out)
    with pytest.raises(fehip.InvalidArgument):
        fehip.elem_restriction_apply(r, fehip.TRANSPOSE, fehip.DeviceArray(3), fehip.DeviceArray(3))
    with pytest.raises(fehip.InvalidArgument):
        fehip.elem_restriction_apply(r, fehip.TRANSPOSE, fehip.DeviceArray(4), out)


def test_fp32_restriction(gpu):
    ne, es, nc, cs, l_size, off = box_restriction_args(5, 4, 3, 3, 2, False)
    r = fehip.elem_restriction_create(ne, es, nc, cs, l_size, off, fehip.SCATTER_DETERMINISTIC)
    l_in = wiggly(l_size, 0.4).astype(np.float32)
    d_l, d_e = fehip.DeviceArray.from_host(l_in), fehip.DeviceArray(nc * es * ne, np.float32)
    fehip.elem_restriction_apply(r, fehip.NO_TRANSPOSE, d_l, d_e)
    e_cpu = oracle().restriction_apply(ne, es, nc, cs, l_size, off, 0, l_in.astype(np.float64))
    assert (d_e.to_host().astype(np.float64) == e_cpu).all()
