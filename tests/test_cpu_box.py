"""Host side of the box-column algorithm: recognising structured boxes from
the offsets, and the algorithm's index arithmetic (every node written exactly
once, directly or by the shared-node pass; every partial slot written once and
read once; owners in ascending order), over boxes of many shapes, every
compiled column cross-section and several layer counts."""
import numpy as np
import pytest

import fehip
from fehip.meshes import box_mesh, l_shape, permute


@pytest.mark.parametrize("n3,P", [((4, 4, 4), 5), ((3, 5, 2), 3), ((6, 2, 7), 4), ((1, 1, 1), 2), ((2, 3, 4), 8)])
def test_box_detect_recognises_boxes(n3, P):
    off = fehip.box_offsets(*n3, P, 1)
    E = n3[0] * n3[1] * n3[2]
    N = np.prod([n * (P - 1) + 1 for n in n3])
    assert fehip.box_detect(E, P, off) == n3
    assert fehip.box_detect(E, P, off, num_comp=3, comp_stride=N, l_size=3 * N) == n3
    # One component: the comp_stride is not used.
    assert fehip.box_detect(E, P, off, num_comp=1, comp_stride=1, l_size=N) == n3


def test_box_detect_rejects_everything_else():
    n, P = 4, 3
    off = fehip.box_offsets(n, n, n, P, 1)
    E, N = n ** 3, (n * (P - 1) + 1) ** 3
    bad = off.copy()
    bad[len(off) // 2 + 3] += 1  # one offset in the middle
    assert fehip.box_detect(E, P, bad, l_size=N) is None
    assert fehip.box_detect(E, P, fehip.box_offsets(n, n, n, P, 2), l_size=2 * N) is None  # interleaved
    assert fehip.box_detect(E, P, off, num_comp=2, comp_stride=N + 1, l_size=2 * N + 1) is None  # padded stride
    assert fehip.box_detect(E, P, off, l_size=N + 5) is None  # spare L entries
    m = permute(box_mesh(n, P), seed=1)
    assert fehip.box_detect(E, P, m["offsets"]) is None
    m = l_shape(n, P)
    assert fehip.box_detect(m["num_elem"], P, m["offsets"]) is None
    # Elements in another order (same node numbering).
    rows = off.reshape(E, -1)[np.random.default_rng(0).permutation(E)]
    assert fehip.box_detect(E, P, rows.ravel(), l_size=N) is None


# (box, P, elems per layer, most layers per segment (0: automatic)); the
# layer limits that do not divide nz give segments of unequal length
# (64 layers in 3 segments: 21, 21, 22).
CHECKS = [((64, 64, 64), 5, 16, 0), ((64, 64, 64), 5, 8, 0), ((8, 8, 8), 5, 16, 2), ((8, 8, 6), 5, 8, 3),
          ((4, 6, 6), 5, 4, 2), ((4, 6, 6), 5, 4, 6), ((4, 6, 4), 5, 8, 1), ((12, 8, 7), 4, 8, 7), ((6, 4, 10), 3, 4, 5),
          ((4, 4, 1), 5, 16, 0), ((4, 4, 2), 2, 16, 2), ((16, 16, 16), 3, 16, 4), ((2, 2, 9), 8, 4, 3),
          ((64, 64, 64), 5, 16, 22), ((8, 8, 7), 5, 8, 3), ((4, 4, 10), 3, 16, 4), ((4, 2, 11), 4, 8, 2),
          ((2, 2, 13), 8, 4, 5), ((8, 4, 9), 2, 16, 4), ((12, 12, 5), 6, 16, 2)]


@pytest.mark.parametrize("n3,P,elems,layers", CHECKS)
def test_box_column_index_arithmetic(n3, P, elems, layers):
    st = fehip.box_column_check(*n3, P, elems, layers)
    N = np.prod([n * (P - 1) + 1 for n in n3])
    assert st["direct_nodes"] + st["shared_nodes"] == N
    bx, by = {16: (4, 4), 8: (4, 2), 4: (2, 2)}[elems]
    ncols = (n3[0] // bx) * (n3[1] // by)
    assert st["segments"] % ncols == 0
    nsz = st["segments"] // ncols
    lz = st["layers"]
    assert lz == -(-n3[2] // nsz)  # the longest segment: ceil(nz / nsz)
    if layers:
        assert nsz == -(-n3[2] // layers) and lz <= layers
    if st["segments"] > 1:
        assert st["max_owners"] in (2, 4, 8)


def test_box_column_layers_automatic():
    # The check assumes 256 CUs and 32 elements per column layer per CU
    # (32 * 256 / elems workgroup slots), and picks the segment count with the
    # fewest layer steps on the busiest CU, fewest segments on a tie: 64^3
    # with 4x4 columns (256 columns, 512 slots) -> 2 segments of 32 layers; 4x2
    # (512 columns, 1024 slots) and 2x2 (1024, 2048) -> also 32 layers.
    assert fehip.box_column_check(64, 64, 64, 5, 16)["layers"] == 32
    assert fehip.box_column_check(64, 64, 64, 5, 8)["layers"] == 32
    assert fehip.box_column_check(64, 64, 64, 5, 4)["layers"] == 32
    # 48 layers over 512 slots: 2 segments of 24 (not a power-of-two divisor).
    assert fehip.box_column_check(64, 64, 48, 5, 16)["layers"] == 24
    # Too few columns for the target: one layer per segment.
    assert fehip.box_column_check(8, 8, 8, 5, 16)["layers"] == 1
    # P = 2 needs two layers per segment (a segment end must hold 2 node planes).
    assert fehip.box_column_check(8, 8, 6, 2, 16)["layers"] == 2


def test_box_column_check_rejects_bad_sizes():
    with pytest.raises(ValueError):
        fehip.box_column_check(6, 4, 4, 5, 16)  # 4x4 columns do not tile nx = 6
    with pytest.raises(ValueError):
        fehip.box_column_check(4, 4, 4, 5, 12)  # no such cross-section
    with pytest.raises(ValueError):
        fehip.box_column_check(4, 4, 6, 2, 16, 1)  # P = 2 segments of one layer hold 2 node planes
