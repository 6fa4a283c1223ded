"""Diffusion (Poisson) operator on the device: the build_diffusion and
apply_diffusion QFunctions through the generic device operator (gradient basis
both ways, deterministic scatter).

* qdata and K u equal the CPU oracle bit for bit (same operations, same order;
  the oracle composes the steps with a zeroed E-vector for the transposed
  gradient, which the device's Grad-Transpose does by itself).
* At 24^3, P=4: K 1 = 0, symmetry, u^T K u = |a|^2 for u linear in x, and the
  interior patch test, to rounding.
"""
import numpy as np
import pytest

import fehip
from fehip.problems import DiffusionProblem
from helpers import node_coords, rel_err
from oracle_lib import oracle

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n,P,Q,nc", [(3, 2, 3, 1), (2, 3, 4, 1), (2, 4, 5, 1), (2, 5, 6, 1), (2, 3, 4, 2)])
def test_diffusion_bitwise_vs_oracle(gpu, n, P, Q, nc):
    print("diffusion vs oracle", n, P, Q, nc, flush=True)
    p = DiffusionProblem(n, P, Q, nc)
    o = oracle()
    nv = (n + 1) ** 3
    want_q = o.build_diffusion_qdata(3, p.num_elem, 2, Q, 3 * nv, nv, p.mesh_offsets, p.coords)
    assert p.qdata().tobytes() == want_q.tobytes()
    u = fehip.rng_uniform(5, p.l_size)
    d_u, d_v = fehip.DeviceArray.from_host(u), fehip.DeviceArray(p.l_size)
    p.op.apply(d_u, d_v)
    want = o.diffusion_apply(3, p.num_elem, P, Q, nc, p.l_size, p.num_nodes, p.offsets, want_q, u)
    assert d_v.to_host().tobytes() == want.tobytes(), rel_err(d_v.to_host(), want)


def test_diffusion_properties_at_scale(gpu):
    n, P, Q = 24, 4, 5
    p = DiffusionProblem(n, P, Q)
    L = p.l_size

    def K(x):
        d_x, d_y = fehip.DeviceArray.from_host(x), fehip.DeviceArray(L)
        p.op.apply(d_x, d_y)
        return d_y.to_host()

    e = np.zeros(L)
    e[L // 2] = 1.0
    scale = np.abs(K(e)).max()
    assert np.abs(K(np.ones(L))).max() <= 1e-12 * scale
    u, w = fehip.rng_uniform(1, L), fehip.rng_uniform(2, L)
    a, b = w @ K(u), u @ K(w)
    assert abs(a - b) <= 1e-12 * abs(a)
    X = node_coords(n, P, p.coords, fehip.gauss_lobatto(P, weights=False))
    g = np.array([0.3, -1.1, 0.7])
    ul = g @ X
    Ku = K(ul)
    assert abs(ul @ Ku - g @ g) <= 1e-11
    N = n * (P - 1) + 1
    idx = np.arange(N)
    gz, gy, gx = np.meshgrid(idx, idx, idx, indexing="ij")
    interior = ((gx > 0) & (gx < N - 1) & (gy > 0) & (gy < N - 1) & (gz > 0) & (gz < N - 1)).ravel()
    assert np.abs(Ku[interior]).max() <= 1e-11 * scale
