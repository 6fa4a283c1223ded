"""Mass problems on structured jittered hex meshes, set up through the device API.

A problem is: the node restriction (lexicographic offsets on the (n(P-1)+1)^3 grid,
blocked components), the H1 Lagrange basis (GLL nodes, Gauss points), the
trilinear mesh (vertex coordinates, P=2 restriction with 3 blocked components),
qdata = det(J) w built on the device by the build_mass operator
(fe-demo-cpp tests/test-operator.cpp:103-112), and the mass operator
(tests/test-operator.cpp:75-82), which runs the fused per-element kernel.
"""
import numpy as np

from . import (EVAL_GRAD, EVAL_NONE, EVAL_WEIGHT, QF_APPLY_DIFFUSION, QF_BUILD_DIFFUSION, QF_BUILD_MASS,
               SCATTER_DETERMINISTIC, VECTOR_ACTIVE, VECTOR_NONE, VECTOR_PASSIVE, DeviceArray, Operator, QFunction,
               TensorBasis, box_coords, box_offsets, elem_restriction_create, elem_restriction_create_strided,
               operator_create_mass)


class MassProblem:
    def __init__(self, n, P, Q, num_comp=1, jitter=0.1, seed=0, scatter=SCATTER_DETERMINISTIC, stream=None):
        self.n, self.P, self.Q, self.nc = n, P, Q, num_comp
        self.num_elem = E = n ** 3
        self.num_nodes = N = (n * (P - 1) + 1) ** 3
        self.l_size = num_comp * N
        self.offsets = box_offsets(n, n, n, P, 1)
        self.rstr = elem_restriction_create(E, P ** 3, num_comp, N, num_comp * N, self.offsets, scatter)
        self.basis = TensorBasis.create_tensor_H1_lagrange(3, num_comp, P, Q)
        # Geometry and qdata.
        nv = (n + 1) ** 3
        self.coords = box_coords(n, jitter, seed)
        self.mesh_offsets = box_offsets(n, n, n, 2, 1)
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

    # Algorithmic bytes and flops of one apply (FP64 unless s_u/s_q are given).
    def apply_bytes(self, s_u=8, s_v=8, s_q=8):
        E, P, Q = self.num_elem, self.P, self.Q
        return s_u * self.l_size + s_v * self.l_size + s_q * E * Q ** 3 + 4 * E * P ** 3

    def apply_flops(self):
        E, P, Q = self.num_elem, self.P, self.Q
        return E * self.nc * (4 * (P ** 3 * Q + P ** 2 * Q ** 2 + P * Q ** 3) + Q ** 3)


class DiffusionProblem:
    """The diffusion (Poisson) operator on the same jittered hex mesh: qdata =
    w det(J) J^-1 J^-T (6 values per point) from the build_diffusion QFunction,
    and K u = sum_e B_grad^T qdata B_grad u_e through the device operator
    (gradient basis both ways, unfused)."""

    def __init__(self, n, P, Q, num_comp=1, jitter=0.1, seed=0, scatter=SCATTER_DETERMINISTIC, stream=None):
        self.n, self.P, self.Q, self.nc = n, P, Q, num_comp
        self.num_elem = E = n ** 3
        self.num_nodes = N = (n * (P - 1) + 1) ** 3
        self.l_size = num_comp * N
        self.offsets = box_offsets(n, n, n, P, 1)
        self.rstr = elem_restriction_create(E, P ** 3, num_comp, N, num_comp * N, self.offsets, scatter)
        self.basis = TensorBasis.create_tensor_H1_lagrange(3, num_comp, P, Q)
        nv = (n + 1) ** 3
        self.coords = box_coords(n, jitter, seed)
        self.mesh_offsets = box_offsets(n, n, n, 2, 1)
        self.mesh_rstr = elem_restriction_create(E, 8, 3, nv, 3 * nv, self.mesh_offsets)
        self.mesh_basis = TensorBasis.create_tensor_H1_lagrange(3, 3, 2, Q)
        self.q_rstr = elem_restriction_create_strided(E, Q ** 3, 6, 6 * E * Q ** 3)
        qf = QFunction(QF_BUILD_DIFFUSION)
        qf.add_input("dx", 9, EVAL_GRAD)
        qf.add_input("weights", 1, EVAL_WEIGHT)
        qf.add_output("qdata", 6, EVAL_NONE)
        qf.set_context_int(3)
        self.build_op = Operator(qf)
        self.build_op.set_field("dx", self.mesh_rstr, self.mesh_basis, VECTOR_ACTIVE)
        self.build_op.set_field("weights", None, self.mesh_basis, VECTOR_NONE)
        self.build_op.set_field("qdata", self.q_rstr, None, VECTOR_ACTIVE)
        self.d_coords = DeviceArray.from_host(self.coords)
        self.d_qdata = DeviceArray(6 * E * Q ** 3)
        self.build_op.apply(self.d_coords, self.d_qdata, stream)
        aq = QFunction(QF_APPLY_DIFFUSION)
        aq.add_input("du", 3 * num_comp, EVAL_GRAD)
        aq.add_input("qdata", 6, EVAL_NONE)
        aq.add_output("dv", 3 * num_comp, EVAL_GRAD)
        aq.set_context_ints(3, num_comp)
        self.op = Operator(aq)
        self.op.set_field("du", self.rstr, self.basis, VECTOR_ACTIVE)
        self.op.set_field("qdata", self.q_rstr, None, VECTOR_PASSIVE, self.d_qdata)
        self.op.set_field("dv", self.rstr, self.basis, VECTOR_ACTIVE)

    def qdata(self):
        return self.d_qdata.to_host()


def host_box_offsets(nx, ny, nz, P, node_stride=1):
    """Independent numpy statement of the lexicographic box offsets (for tests)."""
    Nx, Ny = nx * (P - 1) + 1, ny * (P - 1) + 1
    ez, ey, ex, iz, iy, ix = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), np.arange(P), np.arange(P),
                                         np.arange(P), indexing="ij")
    gx, gy, gz = ex * (P - 1) + ix, ey * (P - 1) + iy, ez * (P - 1) + iz
    return ((gx + Nx * (gy + Ny * gz)) * node_stride).astype(np.int32).ravel()
