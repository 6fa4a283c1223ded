"""Pins the CPU oracle (oracle/liboracle.so, the restatement of the reference's CPU
path) before it is trusted as the parity checker:

1. bit equality with the reference's own sources compiled here (oracle/_ref), where built;
2. bit equality with the golden fixtures written from that build (tests/golden);
3. the reference's own known answers (tests/test-*.cpp restated);
4. what the mathematics fixes: textbook nodes, brute-force Kronecker products,
   Lagrange product formula, brute-force diagonal, dense solves, volume.
Each check is chosen so that a dropped term, wrong sign/index or transposed operand fails.
"""
import json
import os

import numpy as np
import pytest

from helpers import box_restriction_args, ramp, wiggly
from oracle_lib import OracleError, oracle, reference

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

needs_ref = pytest.mark.skipif(reference() is None, reason="reference sources not built here (oracle/_ref)")


# ---------------------------------------------------------------- 1. vs the reference build


@needs_ref
def test_basis_bitwise_vs_reference():
    o, r = oracle(), reference()
    for mode in (0, 1):
        for P in range(2, 13):
            for Q in range(1, 14):
                if mode == 1 and Q < 2:
                    continue
                a, b = o.basis(3, 2, P, Q, mode), r.basis(3, 2, P, Q, mode)
                for k in a:
                    assert a[k].tobytes() == b[k].tobytes(), (mode, P, Q, k)


@needs_ref
@pytest.mark.parametrize("dim,nc,P,Q", [(1, 1, 4, 5), (2, 2, 3, 3), (3, 3, 3, 5), (3, 1, 5, 6), (3, 2, 4, 2)])
@pytest.mark.parametrize("eval_mode", [1, 2])
@pytest.mark.parametrize("t_mode", [0, 1])
def test_basis_apply_bitwise_vs_reference(dim, nc, P, Q, eval_mode, t_mode):
    ne = 5
    nd = dim if eval_mode == 2 else 1
    in_len = nc * P ** dim * ne if t_mode == 0 else nd * nc * Q ** dim * ne
    out_len = nd * nc * Q ** dim * ne if t_mode == 0 else nc * P ** dim * ne
    u = wiggly(in_len, 0.2)
    a = oracle().basis_apply(dim, nc, P, Q, ne, t_mode, eval_mode, u, out_len)
    b = reference().basis_apply(dim, nc, P, Q, ne, t_mode, eval_mode, u, out_len)
    assert a.tobytes() == b.tobytes()
    if eval_mode == 2 and t_mode == 1:
        # Quirk: Transpose grad accumulates into whatever v held (src/tensor-contract.cpp:115-122,142).
        v0 = wiggly(out_len, 1.3)
        a = oracle().basis_apply(dim, nc, P, Q, ne, t_mode, eval_mode, u, out_len, v_in=v0)
        b = reference().basis_apply(dim, nc, P, Q, ne, t_mode, eval_mode, u, out_len, v_in=v0)
        assert a.tobytes() == b.tobytes()


@needs_ref
def test_weights_restriction_qf_bitwise_vs_reference():
    o, r = oracle(), reference()
    for dim in (1, 2, 3):
        n = 4 * 5 ** dim
        assert o.basis_apply(dim, 1, 3, 5, 4, 0, 3, None, n).tobytes() == \
            r.basis_apply(dim, 1, 3, 5, 4, 0, 3, None, n).tobytes()
    for shape in ((3, 3, 3, 3, 3, False), (3, 2, 2, 3, 2, True)):
        ne, es, nc, cs, l_size, off = box_restriction_args(*shape)
        l_in, e_in = wiggly(l_size, 0.1), wiggly(nc * es * ne, 0.7)
        for t, inp, out in ((0, l_in, None), (1, e_in, np.ones(l_size))):
            assert o.restriction_apply(ne, es, nc, cs, l_size, off, t, inp, out).tobytes() == \
                r.restriction_apply(ne, es, nc, cs, l_size, off, t, inp, out).tobytes()
        assert o.multiplicity(ne, es, nc, cs, l_size, off).tobytes() == \
            r.multiplicity(ne, es, nc, cs, l_size, off).tobytes()
    rng = np.random.default_rng(0)
    for dim in (1, 2, 3):
        J, w = rng.uniform(-1, 1, dim * dim * 50), rng.uniform(0, 1, 50)
        assert o.qf_build_mass(dim, 50, J, w).tobytes() == r.qf_build_mass(dim, 50, J, w).tobytes()


@needs_ref
@pytest.mark.parametrize("n,P,Q,nc", [(4, 3, 4, 1), (3, 5, 6, 1), (3, 4, 5, 3), (2, 8, 10, 2)])
def test_mass_operator_bitwise_vs_reference(n, P, Q, nc):
    import fehip

    E, nv, N = n ** 3, (n + 1) ** 3, (n * (P - 1) + 1) ** 3
    coords = fehip.box_coords(n, 0.1, 0)
    mo = fehip.box_offsets(n, n, n, 2, 1)
    qa = oracle().build_qdata(3, E, 2, Q, 3 * nv, nv, mo, coords)
    qb = reference().build_qdata(3, E, 2, Q, 3 * nv, nv, mo, coords)
    assert qa.tobytes() == qb.tobytes()
    off = fehip.box_offsets(n, n, n, P, 1)
    u = fehip.rng_uniform(1, nc * N)
    va = oracle().mass(3, E, P, Q, nc, nc * N, N, off, qa).apply(u)
    vb = reference().mass(3, E, P, Q, nc, nc * N, N, off, qb).apply(u)
    assert va.tobytes() == vb.tobytes()


# ---------------------------------------------------------------- 2. vs the golden fixtures


def test_basis_matches_golden():
    g = np.load(os.path.join(GOLD, "basis.npz"))
    seen = 0
    for key in g.files:
        mode, P, Q, name = key.split("_", 3)
        b = oracle().basis(3, 1, int(P[1:]), int(Q[1:]), int(mode[1:]))
        assert b[name].tobytes() == g[key].tobytes(), key
        seen += 1
    assert seen == json.load(open(os.path.join(GOLD, "manifest.json")))["basis"]["count"]


def test_c1_mass_matches_golden():
    import hashlib

    import fehip

    man = json.load(open(os.path.join(GOLD, "manifest.json")))["c1"]
    g = np.load(os.path.join(GOLD, "c1_mass.npz"))
    n, P, Q = 16, 3, 4
    E, nv, N = n ** 3, (n + 1) ** 3, (n * (P - 1) + 1) ** 3
    ins = dict(offsets=fehip.box_offsets(n, n, n, P, 1), mesh_offsets=fehip.box_offsets(n, n, n, 2, 1),
               coords=fehip.box_coords(n, 0.1, 0), u=fehip.rng_uniform(1, N))
    for k, v in ins.items():
        assert hashlib.sha256(v.tobytes()).hexdigest() == man["inputs"][k], k
    qd = oracle().build_qdata(3, E, 2, Q, 3 * nv, nv, ins["mesh_offsets"], ins["coords"])
    assert hashlib.sha256(qd.tobytes()).hexdigest() == man["qdata"]
    assert qd[::97].tobytes() == g["qdata_sample"].tobytes()
    v = oracle().mass(3, E, P, Q, 1, N, N, ins["offsets"], qd).apply(ins["u"])
    assert v.tobytes() == g["v"].tobytes()


# ---------------------------------------------------------------- 3. the reference's known answers


def test_quadrature_known_answers():
    """tests/test-basis.cpp:12-102"""
    o = oracle()
    for Q in range(1, 6):
        x, w = o.gauss(Q)
        assert abs(w.sum() - 2) < 1e-12
        assert (np.diff(x) > 0).all() and x.min() >= -1 and x.max() <= 1
        for k in range(2 * Q):
            exact = 2.0 / (k + 1) if k % 2 == 0 else 0.0
            assert abs(np.sum(w * x ** k) - exact) < 1e-10
    for P in range(2, 6):
        x, w = o.lobatto(P)
        assert abs(x[0] + 1) < 1e-12 and abs(x[-1] - 1) < 1e-12 and (np.diff(x) > 0).all()
        assert abs(w.sum() - 2) < 1e-12
    x, _ = o.lobatto(3, weights=False)
    assert abs(x[1]) < 1e-12
    # textbook nodes
    x, w = o.gauss(2)
    np.testing.assert_allclose(x, [-1 / np.sqrt(3), 1 / np.sqrt(3)], rtol=0, atol=1e-15)
    x, w = o.gauss(3)
    np.testing.assert_allclose(x, [-np.sqrt(0.6), 0, np.sqrt(0.6)], rtol=0, atol=1e-15)
    np.testing.assert_allclose(w, [5 / 9, 8 / 9, 5 / 9], rtol=1e-15)
    # against numpy's Golub-Welsch rule
    for Q in range(1, 16):
        x, w = o.gauss(Q)
        xg, wg = np.polynomial.legendre.leggauss(Q)
        np.testing.assert_allclose(x, xg, atol=2e-15)
        np.testing.assert_allclose(w, wg, rtol=1e-13)


def test_quadrature_middle_slot_quirk():
    """Odd Q: the middle node is written as -x then x (src/basis.cpp:65-66): a tiny
    non-negative value, not -0 or a negative one."""
    for Q in (3, 5, 7, 9):
        x, _ = oracle().gauss(Q)
        assert np.signbit(x[Q // 2]) == False and abs(x[Q // 2]) < 1e-15  # noqa: E712


def test_lagrange_known_answers_and_product_formula():
    """tests/test-basis.cpp:108-206 plus the product formula L_j(x) = prod (x-x_k)/(x_j-x_k)."""
    o = oracle()
    I, G = o.lagrange([-1.0, 0.0, 1.0], [-1.0, 0.0, 1.0])
    np.testing.assert_allclose(I.reshape(3, 3), np.eye(3), atol=1e-10)
    I, G = o.lagrange([-1, -0.3, 0.4, 1], [-1, -0.6, 0, 0.55, 1])
    np.testing.assert_allclose(I.reshape(5, 4).sum(1), 1, atol=1e-10)
    np.testing.assert_allclose(G.reshape(5, 4).sum(1), 0, atol=1e-10)
    xs = np.array([-1, -0.5, 0, 0.5, 1.0])
    I, G = o.lagrange([-1.0, 1.0], xs)
    np.testing.assert_allclose(I.reshape(5, 2), np.stack([(1 - xs) / 2, (1 + xs) / 2], 1), atol=1e-12)
    np.testing.assert_allclose(G.reshape(5, 2), np.tile([-0.5, 0.5], (5, 1)), atol=1e-12)
    b = o.basis(1, 1, 2, 2)
    g = b["grad"].reshape(2, 2)
    K = np.einsum("q,qi,qj->ij", b["q_weight"], g, g)
    np.testing.assert_allclose(K, [[0.5, -0.5], [-0.5, 0.5]], atol=1e-10)
    nodes, _ = o.lobatto(6)
    pts, _ = o.gauss(7)
    I, G = o.lagrange(nodes, pts)
    for m, x in enumerate(pts):
        for j in range(6):
            others = [k for k in range(6) if k != j]
            Lj = np.prod([(x - nodes[k]) / (nodes[j] - nodes[k]) for k in others])
            dLj = sum(np.prod([(x - nodes[k]) / (nodes[j] - nodes[k]) for k in others if k != i]) /
                      (nodes[j] - nodes[i]) for i in others)
            assert abs(I[m * 6 + j] - Lj) < 1e-13 and abs(G[m * 6 + j] - dLj) < 1e-12


def test_contract_known_answers():
    """tests/test-tensor-contract.cpp:17-104"""
    o = oracle()
    kT = np.array([1.0, 2, 3, 4, 5, 6])
    assert list(o.contract(1, 3, 1, 2, kT, 0, 0, [2, -1, 3])) == [9, 21]
    assert list(o.contract(2, 3, 1, 2, kT, 0, 0, [2, -1, 3, 1, 1, 1])) == [9, 21, 6, 15]
    assert list(o.contract(1, 3, 2, 2, kT, 0, 0, [2, 1, -1, 1, 3, 1])) == [9, 6, 21, 15]
    assert list(o.contract(1, 3, 1, 2, kT, 0, 1, [2, -1, 3], [100, 200])) == [109, 221]
    assert list(o.contract(1, 2, 1, 3, kT, 1, 0, [9, 21])) == [93, 123, 153]


@pytest.mark.parametrize("dim,nc,P,Q", [(1, 1, 3, 4), (2, 2, 3, 4), (3, 1, 3, 4), (3, 3, 2, 3), (3, 1, 4, 2)])
def test_interp_grad_equal_kronecker_products(dim, nc, P, Q):
    """Brute force: interp = kron(B,..,B) per component and element (x fastest), grad
    block d = kron with G on axis d; transposes are the matrix transposes."""
    o = oracle()
    b = o.basis(dim, nc, P, Q)
    B, G = b["interp"].reshape(Q, P), b["grad"].reshape(Q, P)
    ne = 3

    def kron_axes(mats):  # mats[d] acts on axis d; axis 0 fastest
        K = np.ones((1, 1))
        for M in mats:
            K = np.kron(M, K)
        return K

    u = wiggly(nc * P ** dim * ne, 0.4)
    U = u.reshape(nc, P ** dim, ne)
    v = o.basis_apply(dim, nc, P, Q, ne, 0, 1, u, nc * Q ** dim * ne).reshape(nc, Q ** dim, ne)
    KB = kron_axes([B] * dim)
    np.testing.assert_allclose(v, np.einsum("qp,cpe->cqe", KB, U), atol=1e-13)
    g = o.basis_apply(dim, nc, P, Q, ne, 0, 2, u, dim * nc * Q ** dim * ne).reshape(dim, nc, Q ** dim, ne)
    for d in range(dim):
        KG = kron_axes([G if a == d else B for a in range(dim)])
        np.testing.assert_allclose(g[d], np.einsum("qp,cpe->cqe", KG, U), atol=1e-12)
    w = wiggly(dim * nc * Q ** dim * ne, 0.9)
    W = w.reshape(dim, nc, Q ** dim, ne)
    vt = o.basis_apply(dim, nc, P, Q, ne, 1, 1, w[: nc * Q ** dim * ne], nc * P ** dim * ne)
    np.testing.assert_allclose(vt.reshape(nc, P ** dim, ne), np.einsum("qp,cqe->cpe", KB, W[0]), atol=1e-13)
    gt = o.basis_apply(dim, nc, P, Q, ne, 1, 2, w, nc * P ** dim * ne).reshape(nc, P ** dim, ne)
    want = sum(np.einsum("qp,cqe->cpe", kron_axes([G if a == d else B for a in range(dim)]), W[d])
               for d in range(dim))
    np.testing.assert_allclose(gt, want, atol=1e-12)


def test_weights_known_answers():
    """tests/test-tensor-contract.cpp:577-647"""
    o = oracle()
    for d in (1, 2, 3):
        v = o.basis_apply(d, 3, 3, 5, 2, 0, 3, None, 2 * 5 ** d)
        assert abs(v.sum() - 2 * 2 ** d) < 1e-12
    b = o.basis(3, 1, 3, 4)
    w, x = b["q_weight"], b["q_ref"]
    v = o.basis_apply(3, 1, 3, 4, 2, 0, 3, None, 2 * 64).reshape(64, 2)
    for qz in range(4):
        for qy in range(4):
            for qx in range(4):
                q = (qz * 4 + qy) * 4 + qx
                assert (v[q] == w[qz] * (w[qy] * w[qx])).all()
    v1 = v[:, 0]
    X = np.array([(x[q % 4], x[(q // 4) % 4], x[q // 16]) for q in range(64)])
    assert abs(np.sum(v1 * X[:, 0] ** 2 * X[:, 1] ** 4 * (1 + X[:, 2])) - (2 / 3) * (2 / 5) * 2) < 1e-13
    with pytest.raises(OracleError):
        o.basis_apply(2, 1, 3, 4, 3, 1, 3, None, 48)


def test_restriction_known_answers():
    """tests/test-elem-restriction.cpp hand examples and create validation."""
    o = oracle()
    e = o.restriction_apply(2, 2, 1, 1, 3, [0, 1, 1, 2], 0, [10.0, 20, 30])
    assert list(e) == [10, 20, 20, 30]  # [i][e]: (i0,e0),(i0,e1),(i1,e0),(i1,e1)
    l = o.restriction_apply(2, 2, 1, 1, 3, [0, 1, 1, 2], 1, [1.0, 2, 3, 4], np.zeros(3))
    assert list(l) == [1, 5, 4]
    l = o.restriction_apply(2, 2, 1, 1, 3, [0, 1, 1, 2], 1, [1.0, 2, 3, 4], np.full(3, 10.0))
    assert list(l) == [11, 15, 14]
    with pytest.raises(OracleError):
        o.restriction_apply(2, 2, 1, 1, 3, [0, 1, 1, 3], 0, [0.0, 0, 0])
    with pytest.raises(OracleError):
        o.restriction_apply(2, 2, 2, 2, 3, [0, 1, 1, 2], 0, [0.0, 0, 0])
    with pytest.raises(OracleError):
        o.restriction_apply(0, 2, 1, 1, 3, [], 0, [0.0, 0, 0])
    ne, es, nc = 3, 4, 2
    L = wiggly(ne * es * nc, 0.3)
    e = o.restriction_apply(ne, es, nc, 0, ne * es * nc, None, 0, L, backend=True)
    assert e.tobytes() == L.tobytes()  # backend strides: a copy (tests/test-elem-restriction.cpp:303-315)
    # multiplicity (libCEED t209 pattern)
    ne, es, nc, cs, l_size, off = box_restriction_args(3, 2, 1, 3, 1, False)
    m = o.multiplicity(ne, es, nc, cs, l_size, off)
    counts = np.bincount(off, minlength=l_size)
    assert (m == counts).all()


def test_transpose_map_known_answers():
    """tests/cuda/test-cuda-elem-restriction.cu:87-128"""
    o = oracle()
    lv, to, ti = o.transpose_map(2, 2, 3, [0, 1, 1, 2])
    assert list(lv) == [0, 1, 2] and list(to) == [0, 1, 3, 4] and list(ti) == [0, 1, 2, 3]
    lv, to, ti = o.transpose_map(2, 2, 6, [2, 4, 5, 2])
    assert list(lv) == [2, 4, 5] and list(to) == [0, 2, 3, 4] and list(ti) == [0, 3, 1, 2]
    ne, es, nc, cs, l_size, off = box_restriction_args(3, 2, 2, 3, 1, False)
    lv, to, ti = o.transpose_map(ne, es, l_size, off)
    seen = np.zeros(ne * es, int)
    for r in range(len(lv)):
        rows = ti[to[r]:to[r + 1]]
        assert (np.diff(rows) > 0).all()
        assert (off[rows] == lv[r]).all()
        seen[rows] += 1
    assert (seen == 1).all() and (np.diff(lv) > 0).all()


def test_qfunction_known_answers():
    """tests/test-qfunction.cpp:84-139"""
    o = oracle()
    assert list(o.qf_apply_mass(5, [1.0, 2, 3, 4, 5], [0.5, -1, 2, 0, 10])) == [0.5, -2, 6, 0, 50]
    w = [0.5, 2.0]
    assert list(o.qf_build_mass(1, 2, [3.0, -2.0], w)) == [1.5, -4.0]
    assert list(o.qf_build_mass(2, 2, [2, 1, 1, 2, 0, 3, 3, 4], w)) == [3.0, -4.0]
    assert list(o.qf_build_mass(3, 2, [1, 2, 0, 0, 0, 1, 0, 1, 2, 3, 0, 0, 0, 0, 0, 1, 3, 4], w)) == [3.0, 50.0]
    # cofactor determinant against numpy on random matrices
    rng = np.random.default_rng(3)
    J = rng.uniform(-1, 1, (3, 3, 40))
    qd = o.qf_build_mass(3, 40, J.reshape(-1), np.ones(40))
    np.testing.assert_allclose(qd, np.linalg.det(np.moveaxis(J, 2, 0)), rtol=1e-12, atol=1e-15)


def test_operator_known_answers():
    """tests/test-operator.cpp:256-329: 1D mass matrix, build qdata = w h/2, 2D area 5, symmetry."""
    o = oracle()
    b = o.basis(1, 1, 2, 3)
    qd = np.empty(6)
    for q in range(3):
        qd[2 * q], qd[2 * q + 1] = b["q_weight"][q] * 0.5, b["q_weight"][q] * 0.5
    m = o.mass(1, 2, 2, 3, 1, 3, 1, [0, 1, 1, 2], qd)
    np.testing.assert_allclose(m.apply([1.0, 1, 1]), [0.5, 1, 0.5], rtol=1e-14)
    np.testing.assert_allclose(m.apply([0.0, 1, 2]), [1 / 6, 1, 5 / 6], rtol=1e-14)
    q1 = o.build_qdata(1, 2, 2, 3, 3, 3, [0, 1, 1, 2], [0.0, 1, 3])
    np.testing.assert_allclose(q1, np.repeat(b["q_weight"], 2) * np.tile([0.5, 1.0], 3), rtol=1e-14)
    coords = [0, 1, 3, 0.5, 1, 2.5, 0, 0, 0, 2, 2, 2]
    off = [0, 1, 3, 4, 1, 2, 4, 5]
    q2 = o.build_qdata(2, 2, 2, 3, 12, 6, off, coords)
    assert abs(q2.sum() - 5) < 1e-13
    m2 = o.mass(2, 2, 2, 3, 1, 6, 1, off, q2)
    assert abs(m2.apply(np.ones(6)).sum() - 5) < 1e-13
    u, w = ramp(6, 1.0), ramp(6, -0.7)
    assert abs(w @ m2.apply(u) - u @ m2.apply(w)) < 1e-12


# ---------------------------------------------------------------- 4. closed forms and brute force


def test_mass_volume_diagonal_and_cg():
    """On a jittered 3^3, P=3, Q=4 mesh: 1^T M 1 = volume = 1 (boundary vertices are
    fixed); the assembled diagonal equals e_i^T M e_i by brute force; the oracle CG
    solution equals a dense solve."""
    import fehip

    n, P, Q = 3, 3, 4
    E, nv, N = n ** 3, (n + 1) ** 3, (n * (P - 1) + 1) ** 3
    qd = oracle().build_qdata(3, E, 2, Q, 3 * nv, nv, fehip.box_offsets(n, n, n, 2, 1), fehip.box_coords(n, 0.1, 0))
    assert qd.min() > 0 and abs(qd.sum() - 1) < 1e-14
    m = oracle().mass(3, E, P, Q, 1, N, N, fehip.box_offsets(n, n, n, P, 1), qd)
    M = np.stack([m.apply(np.eye(N)[i]) for i in range(N)], 1)
    assert abs(M.sum() - 1) < 1e-13
    np.testing.assert_allclose(M, M.T, atol=1e-15)
    np.testing.assert_allclose(m.diagonal(), np.diag(M), rtol=1e-13)
    b = M @ fehip.rng_uniform(2, N)
    x, it, rel = m.cg(b, m.diagonal(), rtol=1e-12)
    assert rel <= 1e-12
    np.testing.assert_allclose(x, np.linalg.solve(M, b), rtol=0, atol=1e-9)


def test_vector_mass_is_componentwise_scalar_mass():
    import fehip

    n, P, Q, nc = 3, 3, 4, 3
    E, nv, N = n ** 3, (n + 1) ** 3, (n * (P - 1) + 1) ** 3
    qd = oracle().build_qdata(3, E, 2, Q, 3 * nv, nv, fehip.box_offsets(n, n, n, 2, 1), fehip.box_coords(n, 0.1, 0))
    off = fehip.box_offsets(n, n, n, P, 1)
    mv = oracle().mass(3, E, P, Q, nc, nc * N, N, off, qd)
    ms = oracle().mass(3, E, P, Q, 1, N, N, off, qd)
    u = fehip.rng_uniform(4, nc * N)
    v = mv.apply(u)
    for c in range(nc):
        assert v[c * N:(c + 1) * N].tobytes() == ms.apply(u[c * N:(c + 1) * N]).tobytes()
