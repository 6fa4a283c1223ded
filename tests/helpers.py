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


def node_coords(n, P, coords, gll):
    """Physical coordinates [3][N^3] (node gx + N*(gy + N*gz)) of the order-P nodes
    of the jittered n^3 box: the trilinear map of each element's 8 vertices
    (coords [3][(n+1)^3], x fastest) evaluated at the GLL points `gll`."""
    c = np.asarray(coords).reshape(3, n + 1, n + 1, n + 1)  # [axis][z][y][x]
    N = n * (P - 1) + 1
    g = np.arange(N)
    e = np.minimum(g // (P - 1), n - 1)
    t = (np.asarray(gll)[g - e * (P - 1)] + 1) / 2
    out = np.zeros((3, N, N, N))
    for dz in (0, 1):
        wz = (t if dz else 1 - t)[:, None, None]
        for dy in (0, 1):
            wy = (t if dy else 1 - t)[None, :, None]
            for dx in (0, 1):
                wx = (t if dx else 1 - t)[None, None, :]
                out += (wz * wy * wx)[None] * c[:, (e + dz)[:, None, None], (e + dy)[None, :, None],
                                                (e + dx)[None, None, :]]
    return out.reshape(3, -1)
