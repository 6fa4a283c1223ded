"""Meshes beyond the structured box: any element order, any node numbering,
irregular valence, and their files (FEMHMSH1, fem_hip_mesh_read/_write).

A mesh is a dict: dim, num_elem, P, num_nodes, num_vertices, offsets
[E*P^dim] (node restriction, local node i = (iz*P + iy)*P + ix), vertex_offsets
[E*2^dim] (the element's vertices in the same x-fastest order) and coords
[dim][num_vertices].

* box_mesh: the jittered n^3 box of the benchmarks (lexicographic numbering).
* carve: keep a subset of elements and number the remaining nodes and vertices
  compactly; vertices on the new boundary get their unjittered positions, so the
  domain's volume stays exact (an L-shaped domain: 3/4 of the cube).
* permute: shuffle element order, node numbering and vertex numbering.
* MeshProblem: the mass operator of a mesh (qdata through build_mass on the
  device), like problems.MassProblem for the box.
"""
import numpy as np

from . import (EVAL_GRAD, EVAL_NONE, EVAL_WEIGHT, QF_BUILD_MASS, SCATTER_DETERMINISTIC, VECTOR_ACTIVE, VECTOR_NONE,
               DeviceArray, Operator, QFunction, TensorBasis, box_coords, box_offsets, elem_restriction_create,
               elem_restriction_create_strided, operator_create_mass)


def box_mesh(n, P, jitter=0.1, seed=0):
    coords = box_coords(n, jitter, seed)
    ideal = box_coords(n, 0.0, seed)
    return {"dim": 3, "num_elem": n ** 3, "P": P, "num_nodes": (n * (P - 1) + 1) ** 3, "num_vertices": (n + 1) ** 3,
            "offsets": box_offsets(n, n, n, P, 1), "vertex_offsets": box_offsets(n, n, n, 2, 1), "coords": coords,
            "ideal_coords": ideal}


def carve(mesh, keep):
    """Elements where keep[e]; nodes and vertices renumbered in order of first use.
    Vertices that lose an incident element (the new boundary) take their
    unjittered positions when the mesh carries them ("ideal_coords")."""
    keep = np.asarray(keep, bool)
    dim, P = mesh["dim"], mesh["P"]
    es, vs = P ** dim, 2 ** dim
    off = mesh["offsets"].reshape(-1, es)[keep]
    voff = mesh["vertex_offsets"].reshape(-1, vs)[keep]
    nodes, node_new = np.unique(off.ravel(), return_inverse=True)
    verts, vert_new = np.unique(voff.ravel(), return_inverse=True)
    coords = mesh["coords"].reshape(dim, -1)[:, verts].copy()
    if "ideal_coords" in mesh:
        ideal = mesh["ideal_coords"].reshape(dim, -1)
        before = np.bincount(mesh["vertex_offsets"], minlength=mesh["num_vertices"])
        after = np.bincount(voff.ravel(), minlength=mesh["num_vertices"])
        moved = (after < before)[verts]
        coords[:, moved] = ideal[:, verts][:, moved]
    out = {"dim": dim, "num_elem": int(keep.sum()), "P": P, "num_nodes": len(nodes), "num_vertices": len(verts),
           "offsets": node_new.astype(np.int32), "vertex_offsets": vert_new.astype(np.int32),
           "coords": coords.ravel()}
    if "ideal_coords" in mesh:
        out["ideal_coords"] = mesh["ideal_coords"].reshape(dim, -1)[:, verts].ravel()
    return out


def l_shape(n, P, jitter=0.1, seed=0):
    """The n^3 box without the elements with x and y in the upper half: volume 3/4.
    Nodes along the re-entrant edge touch 6 elements instead of 8."""
    m = box_mesh(n, P, jitter, seed)
    e = np.arange(n ** 3)
    ex, ey = e % n, (e // n) % n
    return carve(m, ~((ex >= n // 2) & (ey >= n // 2)))


def permute(mesh, seed=0):
    """The same mesh with shuffled elements, node numbers and vertex numbers."""
    rng = np.random.default_rng(seed)
    dim, P = mesh["dim"], mesh["P"]
    es, vs = P ** dim, 2 ** dim
    pe = rng.permutation(mesh["num_elem"])
    pn = rng.permutation(mesh["num_nodes"]).astype(np.int32)
    pv = rng.permutation(mesh["num_vertices"]).astype(np.int32)
    coords = np.empty_like(mesh["coords"].reshape(dim, -1))
    coords[:, pv] = mesh["coords"].reshape(dim, -1)
    out = dict(mesh)
    out["offsets"] = pn[mesh["offsets"].reshape(-1, es)[pe]].ravel()
    out["vertex_offsets"] = pv[mesh["vertex_offsets"].reshape(-1, vs)[pe]].ravel()
    out["coords"] = coords.ravel()
    if "ideal_coords" in mesh:
        ideal = np.empty_like(coords)
        ideal[:, pv] = mesh["ideal_coords"].reshape(dim, -1)
        out["ideal_coords"] = ideal.ravel()
    out["element_perm"], out["node_perm"] = pe, pn
    return out


class MeshProblem:
    """Mass operator of a 3D mesh dict (or FEMHMSH1 file), num_comp blocked components."""

    def __init__(self, mesh, Q, num_comp=1, scatter=SCATTER_DETERMINISTIC, stream=None):
        if isinstance(mesh, (str, bytes)):
            from . import mesh_read
            mesh = mesh_read(mesh)
        assert mesh["dim"] == 3
        self.mesh = mesh
        E, P, N, nv = mesh["num_elem"], mesh["P"], mesh["num_nodes"], mesh["num_vertices"]
        self.num_elem, self.P, self.Q, self.nc, self.num_nodes = E, P, Q, num_comp, N
        self.l_size = num_comp * N
        self.offsets = np.ascontiguousarray(mesh["offsets"], np.int32)
        self.rstr = elem_restriction_create(E, P ** 3, num_comp, N, num_comp * N, self.offsets, scatter)
        self.basis = TensorBasis.create_tensor_H1_lagrange(3, num_comp, P, Q)
        self.mesh_offsets = np.ascontiguousarray(mesh["vertex_offsets"], np.int32)
        self.coords = np.ascontiguousarray(mesh["coords"], np.float64)
        self.mesh_rstr = elem_restriction_create(E, 8, 3, nv, 3 * nv, self.mesh_offsets)
        self.mesh_basis = TensorBasis.create_tensor_H1_lagrange(3, 3, 2, Q)
        self.q_rstr = elem_restriction_create_strided(E, Q ** 3, 1, E * Q ** 3)
        qf = QFunction(QF_BUILD_MASS)
        qf.add_input("dx", 9, EVAL_GRAD)
        qf.add_input("weights", 1, EVAL_WEIGHT)
        qf.add_output("qdata", 1, EVAL_NONE)
        qf.set_context_int(3)
        self.build_op = Operator(qf)
        self.build_op.set_field("dx", self.mesh_rstr, self.mesh_basis, VECTOR_ACTIVE)
        self.build_op.set_field("weights", None, self.mesh_basis, VECTOR_NONE)
        self.build_op.set_field("qdata", self.q_rstr, None, VECTOR_ACTIVE)
        self.d_coords = DeviceArray.from_host(self.coords)
        self.d_qdata = DeviceArray(E * Q ** 3)
        self.build_op.apply(self.d_coords, self.d_qdata, stream)
        self.op = operator_create_mass(self.rstr, self.basis, self.d_qdata)

    def qdata(self):
        return self.d_qdata.to_host()
