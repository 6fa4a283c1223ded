"""Shared inputs for the tests (restatements of the reference test helpers)."""
import numpy as np


def wiggly(n, phase):
    """Values whose sums round differently by order (tests/cuda/test-cuda-elem-restriction.cu:45-49)."""
    i = np.arange(n, dtype=np.float64)
    return np.sin(0.37 * i + phase) + 1e-3 * np.cos(11.0 * i)


def box_restriction_args(nx, ny, nz, P, num_comp, interleaved):
    """Arguments of box_restriction (tests/cuda/test-cuda-elem-restriction.cu:24-42):
    (num_elem, elem_size, num_comp, comp_stride, l_size, offsets)."""
    Nx, Ny, Nz = nx * (P - 1) + 1, ny * (P - 1) + 1, nz * (P - 1) + 1
    num_nodes = Nx * Ny * Nz
    node_stride = num_comp if interleaved else 1
    comp_stride = 1 if interleaved else num_nodes
    ez, ey, ex, iz, iy, ix = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), np.arange(P), np.arange(P),
                                         np.arange(P), indexing="ij")
    gx, gy, gz = ex * (P - 1) + ix, ey * (P - 1) + iy, ez * (P - 1) + iz
    offsets = ((gx + Nx * (gy + Ny * gz)) * node_stride).astype(np.int32).ravel()
    return nx * ny * nz, P ** 3, num_comp, comp_stride, num_nodes * num_comp, offsets


def ramp(n, scale):
    """tests/test-operator.cpp:46-50"""
    i = np.arange(n)
    return scale * (0.37 * i - 1.3 + 0.01 * (i % 7) * (i % 5))


def rel_err(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))
