"""fem_hip_operator_apply_noalloc: the FP64 fused apply with its per-apply
scratch in the library's static device array (no allocation), checked through
the C ABI against fem_hip_operator_apply on the same operator: bit for bit
(both run the same deterministic kernels), on two calls in a row, the second
with a smaller input, for both supported algorithms; and its refusals."""
import numpy as np
import pytest

import fehip
from fehip.problems import MassProblem

pytestmark = pytest.mark.gpu


def _pair(p, alg, seed):
    """(ordinary apply, no-alloc apply) of u ~ U[-1,1] on p with algorithm alg."""
    p.op.set_fused_options(fehip.SCATTER_DETERMINISTIC, exact=0)
    p.op.set_fused_algorithm(alg)
    assert p.op.fused_choice()[0] == alg
    d_u = fehip.DeviceArray.from_host(fehip.rng_uniform(seed, p.l_size))
    d_ref = fehip.DeviceArray.from_host(np.full(p.l_size, np.nan))
    d_na = fehip.DeviceArray.from_host(np.full(p.l_size, np.nan))
    p.op.apply(d_u, d_ref)  # the existing entry point (also makes the operator's state)
    p.op.apply_noalloc(d_u, d_na)
    return d_ref.to_host(), d_na.to_host()


@pytest.mark.parametrize("alg", [fehip.FUSED_BOX_COLUMN, fehip.FUSED_ELEM_TMAP])
def test_noalloc_equals_apply_two_calls_second_smaller(gpu, alg):
    print("apply_noalloc, alg", alg, flush=True)
    cap = fehip.lib().fem_hip_noalloc_scratch_bytes()
    assert cap == 64 << 20
    big = MassProblem(16, 5, 6, 1, scatter=fehip.SCATTER_DETERMINISTIC)
    small = MassProblem(8, 3, 4, 2, scatter=fehip.SCATTER_DETERMINISTIC)  # 4x2 columns tile 8^3
    ref1, na1 = _pair(big, alg, 1)
    ref2, na2 = _pair(small, alg, 2)  # reuses the same array, left holding the first call's data
    assert na1.tobytes() == ref1.tobytes()
    assert na2.tobytes() == ref2.tobytes()
    assert not np.isnan(na2).any()
    # Back to the larger one: the array's contents from the smaller call do not leak in.
    d_u = fehip.DeviceArray.from_host(fehip.rng_uniform(1, big.l_size))
    d_v = fehip.DeviceArray(big.l_size)
    big.op.apply_noalloc(d_u, d_v)
    assert d_v.to_host().tobytes() == ref1.tobytes()


def test_noalloc_refusals(gpu):
    p = MassProblem(4, 3, 4, 1, scatter=fehip.SCATTER_DETERMINISTIC)
    d_u, d_v = fehip.DeviceArray(p.l_size), fehip.DeviceArray(p.l_size)
    d_u.zero()
    # No device state yet: the first ordinary apply makes it.
    with pytest.raises(fehip.FemHipError) as e:
        p.op.apply_noalloc(d_u, d_v)
    assert e.value.code == fehip.ERR_UNSUPPORTED
    # An algorithm whose apply needs more than the scratch (block plans).
    p.op.set_fused_algorithm(fehip.FUSED_BLOCK_PARTIAL)
    p.op.apply(d_u, d_v)
    with pytest.raises(fehip.FemHipError) as e:
        p.op.apply_noalloc(d_u, d_v)
    assert e.value.code == fehip.ERR_UNSUPPORTED
    # Wrong sizes.
    p.op.set_fused_algorithm(fehip.FUSED_ELEM_TMAP)
    with pytest.raises(fehip.InvalidArgument):
        p.op.apply_noalloc(fehip.DeviceArray(p.l_size - 1), d_v)
    # More scratch than the array holds: element outputs of 85,184 P=5 hexes (85 MB).
    q = MassProblem(44, 5, 6, 1, scatter=fehip.SCATTER_DETERMINISTIC)
    q.op.set_fused_algorithm(fehip.FUSED_ELEM_TMAP)
    d_u2, d_v2 = fehip.DeviceArray(q.l_size), fehip.DeviceArray(q.l_size)
    d_u2.zero()
    q.op.apply(d_u2, d_v2)
    with pytest.raises(fehip.InvalidArgument):
        q.op.apply_noalloc(d_u2, d_v2)
