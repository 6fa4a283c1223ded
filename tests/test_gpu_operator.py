"""Device operator_apply (unfused and fused) vs the CPU oracle.

* Unfused path: the reference's known answers (tests/test-operator.cpp:256-340) and
  bit equality with the oracle (same arithmetic, same order).
* Fused mass kernel: with exact arithmetic and the deterministic transpose-map
  scatter it equals the oracle's operator_apply bit for bit; with FMA and/or any
  other scatter algorithm (element atomics, block atomics, block partial sums
  with colour rounds or LDS atomics, the pipelined transpose-map kernel)
  it agrees within 1e-13 relative (the FP64 tolerance), and the deterministic
  algorithms repeat bit for bit.
* The autotuner times every valid algorithm and its choice is the one applied.
* At the 64^3, P=5, Q=6 size, size-independent properties: 1^T M 1 = volume = 1,
  symmetry, deterministic reproducibility, atomic vs deterministic agreement.
"""
import numpy as np
import pytest

import fehip
from fehip.problems import MassProblem
from helpers import ramp, rel_err
from oracle_lib import oracle

pytestmark = pytest.mark.gpu

FP64_RTOL = 1e-13
ALGS = (fehip.FUSED_ELEM_ATOMIC, fehip.FUSED_ELEM_TMAP, fehip.FUSED_BLOCK_ATOMIC, fehip.FUSED_BLOCK_PARTIAL,
        fehip.FUSED_BLOCK_PARTIAL_LDS,
        fehip.FUSED_BOX_COLUMN)
DET_ALGS = (fehip.FUSED_ELEM_TMAP, fehip.FUSED_BLOCK_PARTIAL, fehip.FUSED_BOX_COLUMN)


def usable(op, alg):
    """Whether `alg` is compiled for the operator's shape and fits its mesh (the
    box-column algorithm exists for some shapes and needs a box its columns tile)."""
    if alg != fehip.FUSED_BOX_COLUMN:
        return True
    op.set_fused_algorithm(alg)
    try:
        op.fused_choice()
        return True
    except (ValueError, fehip.FemHipError):
        return False
    finally:
        op.set_fused_algorithm(fehip.FUSED_AUTO)


def mesh1d_qdata(h0, h1):
    b = fehip.TensorBasis.create_tensor_H1_lagrange(1, 1, 2, 3)
    w = np.empty(6)
    for q in range(3):
        w[q * 2 + 0] = b.q_weight_1d[q] * (h0 / 2)
        w[q * 2 + 1] = b.q_weight_1d[q] * (h1 / 2)
    return w


def mesh1d_mass(qdata):
    """tests/test-operator.cpp:57-83"""
    basis = fehip.TensorBasis.create_tensor_H1_lagrange(1, 1, 2, 3)
    rstr = fehip.elem_restriction_create(2, 2, 1, 1, 3, np.array([0, 1, 1, 2]))
    q_rstr = fehip.elem_restriction_create_strided(2, 3, 1, 6)
    qf = fehip.QFunction(fehip.QF_APPLY_MASS)
    qf.add_input("u", 1, fehip.EVAL_INTERP)
    qf.add_input("qdata", 1, fehip.EVAL_NONE)
    qf.add_output("v", 1, fehip.EVAL_INTERP)
    op = fehip.Operator(qf)
    op.set_field("u", rstr, basis, fehip.VECTOR_ACTIVE)
    op.set_field("qdata", q_rstr, None, fehip.VECTOR_PASSIVE, qdata)
    op.set_field("v", rstr, basis, fehip.VECTOR_ACTIVE)
    return op


def apply(op, u, out_len):
    d_u, d_v = fehip.DeviceArray.from_host(np.asarray(u, np.float64)), fehip.DeviceArray(out_len)
    op.apply(d_u, d_v)
    return d_v.to_host()


def test_mesh1d_mass_known_answers(gpu):
    print("Mesh1D mass", flush=True)
    op = mesh1d_mass(mesh1d_qdata(1, 1))
    info = op.info()
    assert (info["num_elem"], info["num_qpts"], info["active_in_l_size"], info["active_out_l_size"]) == (2, 3, 3, 3)
    v = apply(op, [1, 1, 1], 3)
    np.testing.assert_allclose(v, [0.5, 1.0, 0.5], rtol=1e-14)
    v = apply(op, [0, 1, 2], 3)
    np.testing.assert_allclose(v, [1 / 6, 1.0, 5 / 6], rtol=1e-14)
    orc = oracle().mass(1, 2, 2, 3, 1, 3, 1, np.array([0, 1, 1, 2]), mesh1d_qdata(1, 2))
    op2 = mesh1d_mass(mesh1d_qdata(1, 2))
    u = np.array([0.3, -1.2, 2.5])
    assert apply(op2, u, 3).tobytes() == orc.apply(u).tobytes()


def test_operator_overwrites_output(gpu):
    """tests/test-operator.cpp:331-340: out is zeroed, not accumulated into."""
    op = mesh1d_mass(mesh1d_qdata(1, 1))
    d_u = fehip.DeviceArray.from_host(np.ones(3))
    d_v = fehip.DeviceArray.from_host(np.full(3, 99.0))
    op.apply(d_u, d_v)
    np.testing.assert_allclose(d_v.to_host(), [0.5, 1.0, 0.5], rtol=1e-14)
    with pytest.raises(fehip.InvalidArgument):
        op.apply(fehip.DeviceArray(2), d_v)


def test_operator_field_validation(gpu):
    """tests/test-operator.cpp:147-250"""
    basis = fehip.TensorBasis.create_tensor_H1_lagrange(1, 1, 2, 3)
    rstr = fehip.elem_restriction_create(2, 2, 1, 1, 3, np.array([0, 1, 1, 2]))
    q_rstr = fehip.elem_restriction_create_strided(2, 3, 1, 6)
    qf = fehip.QFunction(fehip.QF_APPLY_MASS)
    qf.add_input("u", 1, fehip.EVAL_INTERP)
    qf.add_input("qdata", 1, fehip.EVAL_NONE)
    qf.add_output("v", 1, fehip.EVAL_INTERP)
    op = fehip.Operator(qf)
    with pytest.raises(fehip.InvalidArgument):
        op.set_field("w", rstr, basis, fehip.VECTOR_ACTIVE)
    op.set_field("u", rstr, basis, fehip.VECTOR_ACTIVE)
    with pytest.raises(fehip.InvalidArgument):
        op.set_field("u", rstr, basis, fehip.VECTOR_ACTIVE)
    with pytest.raises(fehip.InvalidArgument):
        op.check_ready()
    with pytest.raises(fehip.InvalidArgument):
        op.set_field("qdata", q_rstr, basis, fehip.VECTOR_PASSIVE, np.zeros(6))
    with pytest.raises(fehip.InvalidArgument):
        op.set_field("qdata", q_rstr, None, fehip.VECTOR_PASSIVE, np.zeros(5))
    with pytest.raises(fehip.InvalidArgument):
        op.set_field("v", rstr, basis, fehip.VECTOR_PASSIVE, np.zeros(3))
    three = fehip.elem_restriction_create_strided(3, 3, 1, 9)
    with pytest.raises(fehip.InvalidArgument):
        op.set_field("qdata", three, None, fehip.VECTOR_PASSIVE, np.zeros(9))
    two_comp = fehip.TensorBasis.create_tensor_H1_lagrange(1, 2, 2, 3)
    with pytest.raises(fehip.InvalidArgument):
        op.set_field("v", rstr, two_comp, fehip.VECTOR_ACTIVE)


def test_mesh2d_build_and_mass(gpu):
    """Distorted 2D quads of area 5 (tests/test-operator.cpp:90-121,288-329), bitwise vs the oracle."""
    print("Mesh2D build + mass", flush=True)
    coords = np.array([0, 1, 3, 0.5, 1, 2.5, 0, 0, 0, 2, 2, 2], np.float64)
    offsets = np.array([0, 1, 3, 4, 1, 2, 4, 5])
    mesh_basis = fehip.TensorBasis.create_tensor_H1_lagrange(2, 2, 2, 3)
    sol_basis = fehip.TensorBasis.create_tensor_H1_lagrange(2, 1, 2, 3)
    coord_rstr = fehip.elem_restriction_create(2, 4, 2, 6, 12, offsets)
    sol_rstr = fehip.elem_restriction_create(2, 4, 1, 1, 6, offsets)
    q_rstr = fehip.elem_restriction_create_strided(2, 9, 1, 18)
    qf = fehip.QFunction(fehip.QF_BUILD_MASS)
    qf.add_input("dx", 4, fehip.EVAL_GRAD)
    qf.add_input("weights", 1, fehip.EVAL_WEIGHT)
    qf.add_output("qdata", 1, fehip.EVAL_NONE)
    qf.set_context_int(2)
    bop = fehip.Operator(qf)
    bop.set_field("dx", coord_rstr, mesh_basis, fehip.VECTOR_ACTIVE)
    bop.set_field("weights", None, mesh_basis, fehip.VECTOR_NONE)
    bop.set_field("qdata", q_rstr, None, fehip.VECTOR_ACTIVE)
    qdata = apply(bop, coords, 18)
    assert qdata.tobytes() == oracle().build_qdata(2, 2, 2, 3, 12, 6, offsets, coords).tobytes()
    assert abs(qdata.sum() - 5.0) < 1e-13
    mqf = fehip.QFunction(fehip.QF_APPLY_MASS)
    mqf.add_input("u", 1, fehip.EVAL_INTERP)
    mqf.add_input("qdata", 1, fehip.EVAL_NONE)
    mqf.add_output("v", 1, fehip.EVAL_INTERP)
    op = fehip.Operator(mqf)
    op.set_field("u", sol_rstr, sol_basis, fehip.VECTOR_ACTIVE)
    op.set_field("qdata", q_rstr, None, fehip.VECTOR_PASSIVE, qdata)
    op.set_field("v", sol_rstr, sol_basis, fehip.VECTOR_ACTIVE)
    assert abs(apply(op, np.ones(6), 6).sum() - 5.0) < 1e-13
    u, w = ramp(6, 1.0), ramp(6, -0.7)
    Mu, Mw = apply(op, u, 6), apply(op, w, 6)
    assert abs(w @ Mu - u @ Mw) < 1e-12
    orc = oracle().mass(2, 2, 2, 3, 1, 6, 1, offsets, qdata)
    assert Mu.tobytes() == orc.apply(u).tobytes()


def oracle_problem(p):
    return oracle().mass(3, p.num_elem, p.P, p.Q, p.nc, p.l_size, p.num_nodes, p.offsets, p.qdata())


@pytest.mark.parametrize("n,P,Q,nc", [(16, 3, 4, 1), (6, 5, 6, 1), (5, 4, 5, 3), (4, 6, 7, 2), (3, 8, 10, 3),
                                      (4, 2, 3, 1)])
def test_fused_mass_matches_oracle(gpu, n, P, Q, nc):
    print("fused mass vs oracle", n, P, Q, nc, flush=True)
    p = MassProblem(n, P, Q, nc, scatter=fehip.SCATTER_DETERMINISTIC)
    assert p.op.info()["fused"] == 1
    # qdata from the device build operator equals the oracle's bit for bit.
    nv = (n + 1) ** 3
    q_orc = oracle().build_qdata(3, p.num_elem, 2, Q, 3 * nv, nv, p.mesh_offsets, p.coords)
    assert p.qdata().tobytes() == q_orc.tobytes()
    orc = oracle_problem(p)
    u = fehip.rng_uniform(1, p.l_size)
    want = orc.apply(u)
    d_u, d_v = fehip.DeviceArray.from_host(u), fehip.DeviceArray(p.l_size)
    p.op.set_fused_options(fehip.SCATTER_DETERMINISTIC, exact=1)
    p.op.apply(d_u, d_v)
    assert d_v.to_host().tobytes() == want.tobytes(), rel_err(d_v.to_host(), want)
    # Every scatter algorithm, FP64 within 1e-13 and FP32 within single-precision
    # rounding; outputs start as NaN, so an entry the apply fails to write shows.
    d_u32 = fehip.DeviceArray.from_host(u.astype(np.float32))
    nan64, nan32 = np.full(p.l_size, np.nan), np.full(p.l_size, np.nan, np.float32)
    for alg in ALGS:
        det = alg in DET_ALGS
        p.op.set_fused_options(fehip.SCATTER_DETERMINISTIC if det else fehip.SCATTER_ATOMIC, exact=0)
        if not usable(p.op, alg):
            continue
        p.op.set_fused_algorithm(alg)
        assert p.op.fused_choice()[0] == alg
        d_v = fehip.DeviceArray.from_host(nan64)
        p.op.apply(d_u, d_v)
        got = d_v.to_host()
        assert rel_err(got, want) < FP64_RTOL, alg
        if det:
            d_v = fehip.DeviceArray.from_host(nan64)
            p.op.apply(d_u, d_v)
            assert d_v.to_host().tobytes() == got.tobytes(), alg
        d_v32 = fehip.DeviceArray.from_host(nan32)
        p.op.apply(d_u32, d_v32)
        assert rel_err(d_v32.to_host().astype(np.float64), want) < 2e-6, alg
    # The element kernels through elems_per_block (first compiled packing).
    p.op.set_fused_options(fehip.SCATTER_ATOMIC, exact=0)
    p.op.set_fused_algorithm(fehip.FUSED_ELEM_ATOMIC)
    ne = p.op.fused_choice()[1]
    p.op.set_fused_options(fehip.SCATTER_ATOMIC, exact=0, elems_per_block=ne)
    assert p.op.fused_choice() == (fehip.FUSED_ELEM_ATOMIC, ne)
    p.op.apply(d_u, d_v)
    assert rel_err(d_v.to_host(), want) < FP64_RTOL
    # A deterministic scatter refuses the atomic algorithms.
    p.op.set_fused_options(fehip.SCATTER_DETERMINISTIC, exact=0)
    p.op.set_fused_algorithm(fehip.FUSED_BLOCK_ATOMIC)
    with pytest.raises(fehip.InvalidArgument):
        p.op.apply(d_u, d_v)
    p.op.set_fused_algorithm(fehip.FUSED_AUTO)


def test_autotune_records_fastest_valid_choice(gpu):
    print("autotune", flush=True)
    p = MassProblem(8, 4, 5, 1, scatter=fehip.SCATTER_DETERMINISTIC)
    u = fehip.rng_uniform(3, p.l_size)
    want = oracle_problem(p).apply(u)
    d_u, d_v = fehip.DeviceArray.from_host(u), fehip.DeviceArray(p.l_size)
    fehip.tune_table_clear()
    try:
        p.op.set_fused_options(fehip.SCATTER_ATOMIC, exact=0)
        (alg, ne, ms), rows = p.op.autotune(reps=3)
        assert {r[0] for r in rows} == {a for a in ALGS if usable(p.op, a)}
        assert ms == min(r[2] for r in rows) and (alg, ne, ms) in rows
        box = int(usable(p.op, fehip.FUSED_BOX_COLUMN))  # a box row where box-column kernels exist
        assert (4, 5, 1, 0, 0, box, alg, ne) in [t[:8] for t in fehip.tune_table()]
        assert p.op.fused_choice() == (alg, ne)
        p.op.apply(d_u, d_v)
        assert rel_err(d_v.to_host(), want) < FP64_RTOL
        # Deterministic mode times only the deterministic algorithms.
        p.op.set_fused_options(fehip.SCATTER_DETERMINISTIC, exact=0)
        (alg2, ne2, _), rows2 = p.op.autotune(fp32=True, reps=2)
        assert {r[0] for r in rows2} == {a for a in DET_ALGS if usable(p.op, a)}
        assert p.op.fused_choice(fp32=True) == (alg2, ne2)
        # A table row set by hand steers automatic choice.
        fehip.tune_table_set(4, 5, 1, 0, 1, fehip.FUSED_ELEM_TMAP, 23, 0.01, 512)
        assert p.op.fused_choice() == (fehip.FUSED_ELEM_TMAP, 23)
        with pytest.raises(fehip.InvalidArgument):
            fehip.tune_table_set(4, 5, 1, 0, 1, fehip.FUSED_BLOCK_ATOMIC, 16, 0.01, 512)
        # Rows measured on other mesh sizes: the one nearest this mesh (512
        # elements) by ratio steers the choice.
        ne_bp = min(r[1] for r in rows if r[0] == fehip.FUSED_BLOCK_PARTIAL)
        fehip.tune_table_clear()
        fehip.tune_table_set(4, 5, 1, 0, 1, fehip.FUSED_ELEM_TMAP, 23, 0.01, 100_000, box=box)
        fehip.tune_table_set(4, 5, 1, 0, 1, fehip.FUSED_BLOCK_PARTIAL, ne_bp, 0.01, 100, box=box)
        assert p.op.fused_choice() == (fehip.FUSED_BLOCK_PARTIAL, ne_bp)  # 5x below, against 195x above
        fehip.tune_table_set(4, 5, 1, 0, 1, fehip.FUSED_ELEM_TMAP, 23, 0.01, 2048, box=box)
        assert p.op.fused_choice() == (fehip.FUSED_ELEM_TMAP, 23)  # 4x above
    finally:
        fehip.tune_table_clear()


def test_unfused_generic_path_matches_fused(gpu):
    """The same mass operator through the generic path (2D basis is not fused) is
    checked elsewhere; here a 3D shape with no compiled fused kernel (P=3, Q=3)
    runs unfused and must equal the oracle bitwise."""
    p = MassProblem(5, 3, 3, 1)
    assert p.op.info()["fused"] == 0
    u = fehip.rng_uniform(1, p.l_size)
    d_u, d_v = fehip.DeviceArray.from_host(u), fehip.DeviceArray(p.l_size)
    p.op.apply(d_u, d_v)
    want = oracle_problem(p).apply(u)
    # atomic restriction: the scatter order differs, so 1e-13
    assert rel_err(d_v.to_host(), want) < FP64_RTOL


def test_fused_c2_properties(gpu):
    """64^3 hexes, P=5, Q=6 (16.97M DoF): volume, symmetry, reproducibility."""
    print("fused c2 properties", flush=True)
    p = MassProblem(64, 5, 6, 1, scatter=fehip.SCATTER_DETERMINISTIC)
    qd = p.qdata()
    assert qd.min() > 0
    assert abs(qd.sum() - 1.0) < 1e-12
    d_one = fehip.DeviceArray.from_host(np.ones(p.l_size))
    d_v = fehip.DeviceArray(p.l_size)
    p.op.set_fused_options(fehip.SCATTER_DETERMINISTIC, exact=0)
    p.op.apply(d_one, d_v)
    assert abs(d_v.to_host().sum() - 1.0) < 1e-12
    u = fehip.rng_uniform(1, p.l_size)
    w = fehip.rng_uniform(7, p.l_size)
    d_u, d_w = fehip.DeviceArray.from_host(u), fehip.DeviceArray.from_host(w)
    d_Mu, d_Mw = fehip.DeviceArray(p.l_size), fehip.DeviceArray(p.l_size)
    p.op.apply(d_u, d_Mu)
    first = d_Mu.to_host()
    for _ in range(3):
        p.op.apply(d_u, d_Mu)
        assert d_Mu.to_host().tobytes() == first.tobytes()
    p.op.apply(d_w, d_Mw)
    a, b = w @ first, u @ d_Mw.to_host()
    assert abs(a - b) <= 1e-12 * abs(a)
    for alg in ALGS:
        det = alg in DET_ALGS
        p.op.set_fused_options(fehip.SCATTER_DETERMINISTIC if det else fehip.SCATTER_ATOMIC, exact=0)
        p.op.set_fused_algorithm(alg)
        assert p.op.fused_choice()[0] == alg  # the box-column algorithm tiles 64^3 (8 layers per workgroup)
        p.op.apply(d_u, d_Mu)
        assert rel_err(d_Mu.to_host(), first) < FP64_RTOL, alg
    p.op.set_fused_options(fehip.SCATTER_DETERMINISTIC, exact=1)
    p.op.apply(d_u, d_Mu)
    assert rel_err(d_Mu.to_host(), first) < FP64_RTOL


def test_mass_diagonal_matches_oracle(gpu):
    p = MassProblem(6, 3, 4, 2)
    d = fehip.DeviceArray(p.l_size)
    p.op.mass_diagonal(d)
    want = oracle_problem(p).diagonal()
    assert rel_err(d.to_host(), want) < 1e-13


def test_build_mass_with_padded_qdata_l_size(gpu):
    """A backend-strided qdata restriction may have a larger l_size than
    num_elem*Q^3 (src/elem-restriction.cpp:155-181); operator_apply then zeroes
    the whole output first (src/operator.cpp:172), so the tail past the
    element entries is 0, and the element entries equal the oracle's qdata."""
    n, Q = 3, 4
    E, nv = n ** 3, (n + 1) ** 3
    coords = fehip.box_coords(n, 0.1, 3)
    moff = fehip.box_offsets(n, n, n, 2, 1)
    mrstr = fehip.elem_restriction_create(E, 8, 3, nv, 3 * nv, moff)
    mbasis = fehip.TensorBasis.create_tensor_H1_lagrange(3, 3, 2, Q)
    pad = 37
    q_rstr = fehip.elem_restriction_create_strided(E, Q ** 3, 1, E * Q ** 3 + pad)
    qf = fehip.QFunction(fehip.QF_BUILD_MASS)
    qf.add_input("dx", 9, fehip.EVAL_GRAD)
    qf.add_input("weights", 1, fehip.EVAL_WEIGHT)
    qf.add_output("qdata", 1, fehip.EVAL_NONE)
    qf.set_context_int(3)
    bop = fehip.Operator(qf)
    bop.set_field("dx", mrstr, mbasis, fehip.VECTOR_ACTIVE)
    bop.set_field("weights", None, mbasis, fehip.VECTOR_NONE)
    bop.set_field("qdata", q_rstr, None, fehip.VECTOR_ACTIVE)
    d_c = fehip.DeviceArray.from_host(coords)
    d_q = fehip.DeviceArray.from_host(np.full(E * Q ** 3 + pad, np.nan))
    bop.apply(d_c, d_q)
    got = d_q.to_host()
    want = oracle().build_qdata(3, E, 2, Q, 3 * nv, nv, moff, coords)
    assert got[:E * Q ** 3].tobytes() == want.tobytes()
    assert np.all(got[E * Q ** 3:] == 0.0)


def test_retired_algorithms_report_not_compiled(gpu):
    """Algorithm numbers 4 and 6 (retired pipelined forms) select no kernel."""
    p = MassProblem(4, 3, 4, 1, scatter=fehip.SCATTER_DETERMINISTIC)
    for alg in (4, 6):
        p.op.set_fused_options(fehip.SCATTER_ATOMIC, exact=0)
        p.op.set_fused_algorithm(alg)
        with pytest.raises(fehip.InvalidArgument):
            p.op.fused_choice()
    p.op.set_fused_algorithm(fehip.FUSED_AUTO)
    assert p.op.fused_choice()[0] in ALGS
