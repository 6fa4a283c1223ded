"""Diffusion (Poisson) operator oracle, pinned by closed forms (the reference has
no diffusion QFunction; its gradient path, src/tensor-contract.cpp:111-156, is
pinned elsewhere against the reference itself):

* 1D: two linear elements give the tridiagonal [1 -1; -1 2 -1; -1 1]/h; one
  quadratic element on GLL nodes gives [7 -8 1; -8 16 -8; 1 -8 7]/(3L).
* 3D jittered hexes: K 1 = 0; K is symmetric; for u linear in the physical
  coordinates, u^T K u = |a|^2 * volume and (K u) vanishes at interior nodes
  (patch test), up to rounding.
* The restatement and the reference-source build agree bit for bit.
"""
import numpy as np
import pytest

import fehip
from helpers import box_restriction_args, node_coords
from oracle_lib import oracle, reference


def diffusion_1d(orc, coords, P, Q, offsets, ne, l_size):
    mesh_off = np.array([[e, e + 1] for e in range(ne)]).ravel()
    qd = orc.build_diffusion_qdata(1, ne, 2, Q, len(coords), len(coords), mesh_off, coords)
    K = np.stack([orc.diffusion_apply(1, ne, P, Q, 1, l_size, l_size, offsets, qd, np.eye(l_size)[j])
                  for j in range(l_size)], axis=1)
    return K


def test_1d_closed_forms():
    o = oracle()
    K = diffusion_1d(o, np.array([0.0, 1.0, 2.0]), 2, 2, np.array([0, 1, 1, 2]), 2, 3)
    np.testing.assert_allclose(K, [[1, -1, 0], [-1, 2, -1], [0, -1, 1]], atol=1e-15)
    L = 2.0
    K = diffusion_1d(o, np.array([0.0, L]), 3, 3, np.array([0, 1, 2]), 1, 3)
    np.testing.assert_allclose(K, np.array([[7, -8, 1], [-8, 16, -8], [1, -8, 7]]) / (3 * L), atol=1e-14)


def problem_3d(orc, n, P, Q, jitter=0.1, seed=0):
    ne, es, nc, cs, l_size, off = box_restriction_args(n, n, n, P, 1, False)
    coords = fehip.box_coords(n, jitter, seed)
    nv = (n + 1) ** 3
    moff = fehip.box_offsets(n, n, n, 2, 1)
    qd = orc.build_diffusion_qdata(3, ne, 2, Q, 3 * nv, nv, moff, coords)
    apply = lambda u: orc.diffusion_apply(3, ne, P, Q, 1, l_size, cs, off, qd, u)
    return apply, l_size, coords, qd


@pytest.mark.parametrize("n,P,Q", [(3, 2, 3), (2, 3, 4), (2, 4, 5)])
def test_3d_closed_forms(n, P, Q):
    apply, L, coords, qd = problem_3d(oracle(), n, P, Q)
    scale = np.abs(apply(np.eye(L)[L // 2])).max()
    assert np.abs(apply(np.ones(L))).max() <= 1e-13 * scale
    rng = np.random.default_rng(3)
    u, w = rng.uniform(-1, 1, L), rng.uniform(-1, 1, L)
    assert abs(w @ apply(u) - u @ apply(w)) <= 1e-12 * scale * L
    X = node_coords(n, P, coords, fehip.gauss_lobatto(P, weights=False))
    a = np.array([0.3, -1.1, 0.7])
    ul = a @ X
    Ku = apply(ul)
    assert abs(ul @ Ku - a @ a) <= 1e-12
    N = n * (P - 1) + 1
    g = np.arange(N)
    gz, gy, gx = np.meshgrid(g, g, g, indexing="ij")
    interior = ((gx > 0) & (gx < N - 1) & (gy > 0) & (gy < N - 1) & (gz > 0) & (gz < N - 1)).ravel()
    if interior.any():
        assert np.abs(Ku[interior]).max() <= 1e-12 * scale


def test_restatement_equals_reference_build():
    ref = reference()
    if ref is None:
        pytest.skip("reference sources not built here")
    a1, L, _, q1 = problem_3d(oracle(), 2, 3, 4)
    a2, _, _, q2 = problem_3d(ref, 2, 3, 4)
    assert q1.tobytes() == q2.tobytes()
    u = np.random.default_rng(0).uniform(-1, 1, L)
    assert a1(u).tobytes() == a2(u).tobytes()
