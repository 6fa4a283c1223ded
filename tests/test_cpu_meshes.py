"""On-disk meshes (FEMHMSH1) and meshes that are not the lexicographic box.

* write -> read is exact; a flipped byte, a truncated file and a wrong magic
  are rejected.
* On a permuted L-shaped mesh (any element order, node numbering, re-entrant
  edge where nodes touch 6 elements), the library's transpose map equals the
  oracle's and the element-block plan passes its own invariant checks.
* The CPU oracle's mass operator on that mesh: 1^T M 1 = 3/4 (the domain's
  volume, exact because the new boundary vertices are not jittered), and the
  permuted mesh gives the permuted result of the unpermuted one.
"""
import os

import numpy as np
import pytest

import fehip
from fehip.meshes import box_mesh, carve, l_shape, permute
from oracle_lib import oracle

HERE = os.path.dirname(os.path.abspath(__file__))


def test_roundtrip_and_corruption(tmp_path):
    m = permute(l_shape(4, 3), seed=1)
    path = str(tmp_path / "m.femh")
    fehip.mesh_write(path, m)
    r = fehip.mesh_read(path)
    for k in ("dim", "num_elem", "P", "num_nodes", "num_vertices"):
        assert r[k] == m[k]
    assert (r["offsets"] == m["offsets"]).all() and (r["vertex_offsets"] == m["vertex_offsets"]).all()
    assert r["coords"].tobytes() == np.asarray(m["coords"], np.float64).tobytes()
    raw = bytearray(open(path, "rb").read())
    bad = bytearray(raw)
    bad[100] ^= 0x10
    open(path, "wb").write(bad)
    with pytest.raises(fehip.InvalidArgument):
        fehip.mesh_read(path)
    open(path, "wb").write(raw[:-20])
    with pytest.raises(fehip.InvalidArgument):
        fehip.mesh_read(path)
    open(path, "wb").write(b"NOTAMESH" + raw[8:])
    with pytest.raises(fehip.InvalidArgument):
        fehip.mesh_read(path)
    m2 = dict(m)
    m2["offsets"] = m["offsets"].copy()
    m2["offsets"][5] = m["num_nodes"]
    with pytest.raises(fehip.InvalidArgument):
        fehip.mesh_write(path, m2)


def test_irregular_valence_transpose_map_and_plan():
    m = permute(l_shape(4, 3), seed=2)
    E, es, N = m["num_elem"], 27, m["num_nodes"]
    mult = np.bincount(m["offsets"], minlength=N)
    assert set(np.unique(mult)) >= {1, 2, 4, 6, 8}  # 6: nodes on the re-entrant edge
    for a, b in zip(fehip.transpose_map(E, es, N, m["offsets"]), oracle().transpose_map(E, es, N, m["offsets"])):
        assert (a == b).all()
    st = fehip.block_plan_stats(E, es, N, m["offsets"], 16)
    assert st["covers_all"] == 1


def oracle_mass(m, Q=4):
    o = oracle()
    E, P, N, nv = m["num_elem"], m["P"], m["num_nodes"], m["num_vertices"]
    qd = o.build_qdata(3, E, 2, Q, 3 * nv, nv, m["vertex_offsets"], m["coords"])
    return o.mass(3, E, P, Q, 1, N, N, m["offsets"], qd), qd


def test_oracle_mass_on_l_shape():
    m = l_shape(4, 3)
    op, qd = oracle_mass(m)
    assert abs(qd.sum() - 0.75) < 1e-13
    assert abs(op.apply(np.ones(m["num_nodes"])).sum() - 0.75) < 1e-13
    p = permute(m, seed=3)
    op2, _ = oracle_mass(p)
    u = fehip.rng_uniform(1, m["num_nodes"])
    up = np.empty_like(u)
    up[p["node_perm"]] = u
    v, vp = op.apply(u), op2.apply(up)
    np.testing.assert_allclose(vp[p["node_perm"]], v, rtol=0, atol=1e-15 * np.abs(v).max() * 8)


def test_committed_fixture_reads():
    """tests/golden/lshape_n4_p3.femh (written by tests/golden/make_mesh_fixture.py)."""
    path = os.path.join(HERE, "golden", "lshape_n4_p3.femh")
    m = fehip.mesh_read(path)
    assert (m["num_elem"], m["P"]) == (48, 3)
    ref = np.load(os.path.join(HERE, "golden", "lshape_n4_p3.npz"))
    assert (m["offsets"] == ref["offsets"]).all()
    op, qd = oracle_mass(m)
    assert qd.tobytes() == ref["qdata"].tobytes()
    assert op.apply(ref["u"]).tobytes() == ref["v"].tobytes()
