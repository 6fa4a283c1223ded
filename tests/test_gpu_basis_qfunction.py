"""Device basis contractions and QFunctions vs the CPU oracle.

The unfused device kernels perform the reference's arithmetic in its order
(sum from 0 over b ascending, then added into v; no fused multiply-add), so
every comparison here is bit equality (src/tensor-contract.cpp:35-222,
include/qfunctions/mass.hpp:18-51)."""
import numpy as np
import pytest

import fehip
from helpers import wiggly
from oracle_lib import oracle

pytestmark = pytest.mark.gpu

KT = np.array([1, 2, 3, 4, 5, 6], dtype=np.float64)  # tests/test-tensor-contract.cpp:17


def dev_contract(A, B, C, J, t, t_mode, add, u, v0=None):
    d_t, d_u = fehip.DeviceArray.from_host(t), fehip.DeviceArray.from_host(u)
    d_v = fehip.DeviceArray.from_host(np.zeros(A * J * C) if v0 is None else v0)
    fehip.tensor_contract_apply(A, B, C, J, d_t, t_mode, add, d_u, d_v)
    return d_v.to_host()


def test_contract_known_answers(gpu):
    """tests/test-tensor-contract.cpp:19-104"""
    print("contract known answers", flush=True)
    assert list(dev_contract(1, 3, 1, 2, KT, 0, False, np.array([2.0, -1, 3]))) == [9, 21]
    assert list(dev_contract(2, 3, 1, 2, KT, 0, False, np.array([2.0, -1, 3, 1, 1, 1]))) == [9, 21, 6, 15]
    assert list(dev_contract(1, 3, 2, 2, KT, 0, False, np.array([2.0, 1, -1, 1, 3, 1]))) == [9, 6, 21, 15]
    assert list(dev_contract(1, 3, 1, 2, KT, 0, True, np.array([2.0, -1, 3]), np.array([100.0, 200]))) == [109, 221]
    assert list(dev_contract(1, 2, 1, 3, KT, 1, False, np.array([9.0, 21]))) == [93, 123, 153]


@pytest.mark.parametrize("A,B,C,J,t_mode,add", [(3, 4, 7, 5, 0, 0), (2, 5, 33, 6, 1, 1), (1, 10, 100, 12, 0, 1),
                                                (4, 16, 9, 16, 1, 0)])
def test_contract_matches_oracle_bitwise(gpu, A, B, C, J, t_mode, add):
    rng = np.random.default_rng(A * 100 + B)
    t = rng.uniform(-1, 1, J * B)
    u = rng.uniform(-1, 1, A * B * C)
    v0 = rng.uniform(-1, 1, A * J * C)
    got = dev_contract(A, B, C, J, t, t_mode, add, u, v0)
    want = oracle().contract(A, B, C, J, t, t_mode, add, u, v0 if add else None)
    assert got.tobytes() == want.tobytes()


CASES = [(1, 1, 4, 5), (2, 1, 3, 4), (2, 2, 3, 3), (3, 1, 3, 4), (3, 3, 3, 5), (3, 1, 5, 6), (3, 2, 4, 2),
         (3, 1, 8, 10), (3, 3, 2, 3)]


@pytest.mark.parametrize("dim,nc,P,Q", CASES)
@pytest.mark.parametrize("eval_mode", [fehip.EVAL_INTERP, fehip.EVAL_GRAD])
@pytest.mark.parametrize("t_mode", [fehip.NO_TRANSPOSE, fehip.TRANSPOSE])
def test_basis_apply_bitwise(gpu, dim, nc, P, Q, eval_mode, t_mode):
    ne = 7
    b = fehip.TensorBasis.create_tensor_H1_lagrange(dim, nc, P, Q)
    nd = dim if eval_mode == fehip.EVAL_GRAD else 1
    in_len = nc * P ** dim * ne if t_mode == 0 else nd * nc * Q ** dim * ne
    out_len = nd * nc * Q ** dim * ne if t_mode == 0 else nc * P ** dim * ne
    u = wiggly(in_len, 0.3 + dim)
    d_u, d_v = fehip.DeviceArray.from_host(u), fehip.DeviceArray(out_len)
    fehip.tensor_basis_apply(b, ne, t_mode, eval_mode, d_u, d_v)
    want = oracle().basis_apply(dim, nc, P, Q, ne, t_mode, eval_mode, u, out_len)
    assert d_v.to_host().tobytes() == want.tobytes()


@pytest.mark.parametrize("dim,Q", [(1, 4), (2, 3), (3, 4), (3, 6)])
def test_weight_bitwise(gpu, dim, Q):
    ne = 5
    b = fehip.TensorBasis.create_tensor_H1_lagrange(dim, 1, 3, Q)
    d_v = fehip.DeviceArray(Q ** dim * ne)
    fehip.tensor_basis_apply(b, ne, 0, fehip.EVAL_WEIGHT, None, d_v)
    want = oracle().basis_apply(dim, 1, 3, Q, ne, 0, fehip.EVAL_WEIGHT, None, Q ** dim * ne)
    assert d_v.to_host().tobytes() == want.tobytes()
    with pytest.raises(fehip.InvalidArgument):
        fehip.tensor_basis_apply(b, ne, 1, fehip.EVAL_WEIGHT, None, d_v)


def test_adjointness(gpu):
    """<Bu, w> == <u, B^T w> (tests/test-tensor-contract.cpp:543-575)."""
    dim, nc, ne, P, Q = 3, 3, 2, 3, 5
    b = fehip.TensorBasis.create_tensor_H1_lagrange(dim, nc, P, Q)
    rng = np.random.default_rng(33)
    for mode, nd in ((fehip.EVAL_INTERP, 1), (fehip.EVAL_GRAD, dim)):
        u = rng.uniform(-0.5, 0.5, nc * P ** 3 * ne)
        w = rng.uniform(-0.5, 0.5, nd * nc * Q ** 3 * ne)
        d_u, d_Bu = fehip.DeviceArray.from_host(u), fehip.DeviceArray(w.size)
        d_w, d_BTw = fehip.DeviceArray.from_host(w), fehip.DeviceArray(u.size)
        fehip.tensor_basis_apply(b, ne, 0, mode, d_u, d_Bu)
        fehip.tensor_basis_apply(b, ne, 1, mode, d_w, d_BTw)
        assert abs(d_Bu.to_host() @ w - u @ d_BTw.to_host()) < 1e-10


def test_qfunction_apply_mass_known_answer(gpu):
    """tests/test-qfunction.cpp:84-98"""
    qf = fehip.QFunction(fehip.QF_APPLY_MASS)
    qf.add_input("u", 1, fehip.EVAL_INTERP)
    qf.add_input("qdata", 1, fehip.EVAL_NONE)
    qf.add_output("v", 1, fehip.EVAL_INTERP)
    d_u = fehip.DeviceArray.from_host(np.array([1.0, 2, 3, 4, 5]))
    d_q = fehip.DeviceArray.from_host(np.array([0.5, -1, 2, 0, 10]))
    d_v = fehip.DeviceArray(5)
    qf.apply(5, [d_u, d_q], [d_v])
    assert list(d_v.to_host()) == [0.5, -2, 6, 0, 50]
    with pytest.raises(fehip.InvalidArgument):
        qf.apply(5, [d_u], [d_v])
    with pytest.raises(fehip.InvalidArgument):
        qf.apply(4, [d_u, d_q], [d_v])


@pytest.mark.parametrize("dim", [1, 2, 3])
def test_qfunction_build_mass(gpu, dim):
    """tests/test-qfunction.cpp:100-139 determinants, and bit equality on random J."""
    cases = {1: ([3.0, -2.0], 3.0, -2.0), 2: ([2, 1, 1, 2, 0, 3, 3, 4], 6.0, -2.0),
             3: ([1, 2, 0, 0, 0, 1, 0, 1, 2, 3, 0, 0, 0, 0, 0, 1, 3, 4], 6.0, 25.0)}
    J, d0, d1 = cases[dim]
    w = np.array([0.5, 2.0])
    qf = fehip.QFunction(fehip.QF_BUILD_MASS)
    qf.add_input("dx", dim * dim, fehip.EVAL_GRAD)
    qf.add_input("weights", 1, fehip.EVAL_WEIGHT)
    qf.add_output("qdata", 1, fehip.EVAL_NONE)
    qf.set_context_int(dim)
    d_J, d_w, d_q = (fehip.DeviceArray.from_host(np.array(J, np.float64)), fehip.DeviceArray.from_host(w),
                     fehip.DeviceArray(2))
    qf.apply(2, [d_J, d_w], [d_q])
    assert list(d_q.to_host()) == [d0 * 0.5, d1 * 2.0]
    Qn = 1000
    rng = np.random.default_rng(dim)
    J = rng.uniform(-1, 1, dim * dim * Qn)
    w = rng.uniform(0, 1, Qn)
    d_J, d_w, d_q = fehip.DeviceArray.from_host(J), fehip.DeviceArray.from_host(w), fehip.DeviceArray(Qn)
    qf.apply(Qn, [d_J, d_w], [d_q])
    assert d_q.to_host().tobytes() == oracle().qf_build_mass(dim, Qn, J, w).tobytes()


def test_qfunction_field_validation(gpu):
    with pytest.raises(fehip.FemHipError):
        fehip.QFunction(99)  # no device implementation
    qf = fehip.QFunction(fehip.QF_APPLY_MASS)
    with pytest.raises(fehip.InvalidArgument):
        qf.add_input("u", 0, fehip.EVAL_INTERP)
    with pytest.raises(fehip.InvalidArgument):
        qf.add_input("w", 2, fehip.EVAL_WEIGHT)
    with pytest.raises(fehip.InvalidArgument):
        qf.add_output("w", 1, fehip.EVAL_WEIGHT)
    qf.add_input("u", 1, fehip.EVAL_INTERP)
    with pytest.raises(fehip.InvalidArgument):
        qf.add_output("u", 1, fehip.EVAL_INTERP)


def build_op(mesh_rstr, mesh_basis, q_rstr):
    qf = fehip.QFunction(fehip.QF_BUILD_MASS)
    qf.add_input("dx", 9, fehip.EVAL_GRAD)
    qf.add_input("weights", 1, fehip.EVAL_WEIGHT)
    qf.add_output("qdata", 1, fehip.EVAL_NONE)
    qf.set_context_int(3)
    op = fehip.Operator(qf)
    op.set_field("dx", mesh_rstr, mesh_basis, fehip.VECTOR_ACTIVE)
    op.set_field("weights", None, mesh_basis, fehip.VECTOR_NONE)
    op.set_field("qdata", q_rstr, None, fehip.VECTOR_ACTIVE)
    return op


@pytest.mark.parametrize("interleaved,Q,mesh_P", [(False, 2, 2), (True, 4, 2), (True, 7, 2), (False, 3, 3)])
def test_build_mass_operator_matches_oracle_bitwise(gpu, interleaved, Q, mesh_P):
    """qdata through operator_apply of the build_mass operator
    (tests/test-operator.cpp:103-112) equals the oracle's composition bit for
    bit: trilinear geometry (mesh_P = 2) runs the fused geometry kernel, with
    blocked or interleaved coordinates, on a mesh whose elements and vertices
    are permuted; mesh_P = 3 runs the unfused path."""
    from fehip.meshes import l_shape, permute
    print("build_mass operator", interleaved, Q, mesh_P, flush=True)
    m = permute(l_shape(4, 2), seed=9)
    E, nv = m["num_elem"], m["num_vertices"]
    moff, coords = m["vertex_offsets"], m["coords"]  # coords blocked [3][nv]
    if mesh_P == 3:  # quadratic geometry: the midpoints of a 2x refined vertex grid
        n = 3
        moff = fehip.box_offsets(n, n, n, 3, 1)
        nv = (2 * n + 1) ** 3
        g = np.linspace(0.0, 1.0, 2 * n + 1)
        z, y, x = np.meshgrid(g, g, g, indexing="ij")
        coords = np.stack([x.ravel() + 0.01 * np.sin(7 * y.ravel()), y.ravel(), z.ravel() * (1 + 0.1 * x.ravel())])
        coords = coords.ravel()
        E = n ** 3
    es = mesh_P ** 3
    if interleaved:
        off_dev, cs = (np.asarray(moff) * 3).astype(np.int32), 1
        host_coords = coords.reshape(3, nv).T.ravel().copy()
    else:
        off_dev, cs, host_coords = np.asarray(moff, np.int32), nv, coords
    mesh_rstr = fehip.elem_restriction_create(E, es, 3, cs, 3 * nv, off_dev)
    mesh_basis = fehip.TensorBasis.create_tensor_H1_lagrange(3, 3, mesh_P, Q)
    q_rstr = fehip.elem_restriction_create_strided(E, Q ** 3, 1, E * Q ** 3)
    op = build_op(mesh_rstr, mesh_basis, q_rstr)
    d_x = fehip.DeviceArray.from_host(host_coords)
    d_q = fehip.DeviceArray.from_host(np.full(E * Q ** 3, np.nan))
    op.apply(d_x, d_q)
    want = oracle().build_qdata(3, E, mesh_P, Q, 3 * nv, cs, off_dev, host_coords)
    assert d_q.to_host().tobytes() == want.tobytes()
    if mesh_P == 2:
        assert abs(want.sum() - 0.75) < 1e-12  # the L-shape's volume
