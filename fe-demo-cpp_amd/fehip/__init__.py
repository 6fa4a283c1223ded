"""ctypes binding of libfemhip (include/fem_hip.h) for Python callers, tests and bench.py.

Mirrors the reference's interface names (fe-demo-cpp include/cuda/device-array.hpp,
include/cuda/cuda-elem-restriction.hpp, include/basis.hpp, include/tensor-contract.hpp,
include/qfunction.hpp, include/operator.hpp) on device arrays. Errors raise
InvalidArgument (the reference's std::invalid_argument) or FemHipError (its
std::runtime_error). Loading fails loudly when the HIP library is not built: there is
no CPU fallback anywhere in this package.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_HERE), "lib", "libfemhip.so")

SUCCESS, ERR_INVALID_ARGUMENT, ERR_DEVICE, ERR_RUNTIME, ERR_UNSUPPORTED, ERR_OUT_OF_MEMORY = range(6)
NO_TRANSPOSE, TRANSPOSE = 0, 1
EVAL_NONE, EVAL_INTERP, EVAL_GRAD, EVAL_WEIGHT = 0, 1, 2, 3
QUAD_GAUSS, QUAD_GAUSS_LOBATTO = 0, 1
SCATTER_ATOMIC, SCATTER_DETERMINISTIC = 0, 1
FUSED_AUTO, FUSED_ELEM_ATOMIC, FUSED_ELEM_TMAP, FUSED_BLOCK_ATOMIC, FUSED_BLOCK_PARTIAL = -1, 0, 1, 2, 3
FUSED_BLOCK_PARTIAL_LDS, FUSED_BOX_COLUMN = 5, 7  # 4 and 6 are retired
FUSED_NAMES = {0: "elem_atomic", 1: "elem_tmap", 2: "block_atomic", 3: "block_partial", 5: "block_partial_lds",
               7: "box_column"}
VECTOR_ACTIVE, VECTOR_PASSIVE, VECTOR_NONE = 0, 1, 2
QF_BUILD_MASS, QF_APPLY_MASS, QF_APPLY_MASS_VECTOR, QF_BUILD_DIFFUSION, QF_APPLY_DIFFUSION = 1, 2, 3, 4, 5
SOLVE_FP64, SOLVE_MIXED, SOLVE_MIXED_RELIABLE = 0, 1, 2
CG_AUTO, CG_PLAIN, CG_PIPELINED, CG_PLAIN_WIDE = 0, 1, 2, 3  # fem_hip_solve_options.cg_variant


ERR_NOT_READY = 6


class FemHipError(RuntimeError):
    """A failed call (the reference's std::runtime_error)."""

    def __init__(self, code, msg):
        super().__init__(msg)
        self.code = code


class InvalidArgument(ValueError):
    """Bad arguments (the reference's std::invalid_argument)."""

    code = ERR_INVALID_ARGUMENT


class DeviceInfo(ctypes.Structure):
    _fields_ = [
        ("id", ctypes.c_int),
        ("name", ctypes.c_char * 256),
        ("arch", ctypes.c_char * 64),
        ("cu_count", ctypes.c_int),
        ("wavefront_size", ctypes.c_int),
        ("max_threads_per_block", ctypes.c_int),
        ("regs_per_block", ctypes.c_int),
        ("lds_per_block", ctypes.c_size_t),
        ("lds_per_cu", ctypes.c_size_t),
        ("l2_bytes", ctypes.c_size_t),
        ("total_global_mem", ctypes.c_size_t),
        ("free_global_mem", ctypes.c_size_t),
        ("clock_khz", ctypes.c_int),
        ("memory_clock_khz", ctypes.c_int),
        ("memory_bus_width", ctypes.c_int),
    ]


class SolveOptions(ctypes.Structure):
    _fields_ = [
        ("mode", ctypes.c_int),
        ("max_iter", ctypes.c_int),
        ("max_outer", ctypes.c_int),
        ("rtol", ctypes.c_double),
        ("inner_rtol", ctypes.c_double),
        ("cg_variant", ctypes.c_int),
        ("separate_second_pass", ctypes.c_int),
    ]


class SolveStats(ctypes.Structure):
    _fields_ = [
        ("iterations", ctypes.c_int),
        ("outer", ctypes.c_int),
        ("rel_residual", ctypes.c_double),
        ("ms", ctypes.c_double),
    ]


_vp = ctypes.c_void_p
_i = ctypes.c_int
_sz = ctypes.c_size_t
_d = ctypes.c_double
_ip = ctypes.POINTER(ctypes.c_int)
_llp = ctypes.POINTER(ctypes.c_longlong)

# name -> argtypes (every function returns int status unless listed in _RESTYPES)
SIGNATURES = {
    "fem_hip_last_error": [],
    "fem_hip_version": [],
    "fem_hip_device_count": [_ip],
    "fem_hip_device_info_get": [_i, ctypes.POINTER(DeviceInfo)],
    "fem_hip_malloc": [ctypes.POINTER(_vp), _sz],
    "fem_hip_free": [_vp],
    "fem_hip_memcpy_htod": [_vp, _vp, _sz],
    "fem_hip_memcpy_dtoh": [_vp, _vp, _sz],
    "fem_hip_memcpy_dtod": [_vp, _vp, _sz, _vp],
    "fem_hip_memset_zero": [_vp, _sz, _vp],
    "fem_hip_device_synchronize": [],
    "fem_hip_stream_create": [ctypes.POINTER(_vp)],
    "fem_hip_stream_destroy": [_vp],
    "fem_hip_stream_synchronize": [_vp],
    "fem_hip_event_create": [ctypes.POINTER(_vp)],
    "fem_hip_event_destroy": [_vp],
    "fem_hip_event_record": [_vp, _vp],
    "fem_hip_event_synchronize": [_vp],
    "fem_hip_event_elapsed_ms": [_vp, _vp, ctypes.POINTER(ctypes.c_float)],
    "fem_hip_probe_launch": [_i, _i],
    "fem_hip_gauss_legendre": [_i, _vp, _vp],
    "fem_hip_gauss_lobatto": [_i, _vp, _vp],
    "fem_hip_lagrange_basis_matrix": [_i, _vp, _i, _vp, _vp, _vp],
    "fem_hip_restriction_create": [_i, _i, _i, _i, _i, _vp, _i, ctypes.POINTER(_vp)],
    "fem_hip_restriction_create_strided": [_i, _i, _i, _i, _vp, ctypes.POINTER(_vp)],
    "fem_hip_restriction_apply": [_vp, _i, _vp, _sz, _vp, _sz, _vp],
    "fem_hip_restriction_apply_f32": [_vp, _i, _vp, _sz, _vp, _sz, _vp],
    "fem_hip_restriction_apply_arrays": [_vp, _i, _vp, _sz, _vp, _sz, _vp],
    "fem_hip_restriction_get_sizes": [_vp, _llp],
    "fem_hip_restriction_get_transpose_map": [_vp, _vp, _vp, _vp],
    "fem_hip_restriction_destroy": [_vp],
    "fem_hip_transpose_map_build": [_i, _i, _i, _vp, _ip, _vp, _vp, _vp],
    "fem_hip_basis_create_h1_lagrange": [_i, _i, _i, _i, _i, ctypes.POINTER(_vp)],
    "fem_hip_basis_create_tensor": [_i, _i, _i, _i, _vp, _vp, _vp, _vp, ctypes.POINTER(_vp)],
    "fem_hip_basis_get": [_vp, _ip, _vp, _vp, _vp, _vp],
    "fem_hip_basis_apply": [_vp, _i, _i, _i, _vp, _sz, _vp, _sz, _vp],
    "fem_hip_basis_destroy": [_vp],
    "fem_hip_contract": [_i, _i, _i, _i, _vp, _i, _i, _vp, _vp, _vp],
    "fem_hip_qfunction_create": [_i, ctypes.POINTER(_vp)],
    "fem_hip_qfunction_add_input": [_vp, ctypes.c_char_p, _i, _i],
    "fem_hip_qfunction_add_output": [_vp, ctypes.c_char_p, _i, _i],
    "fem_hip_qfunction_set_context": [_vp, _vp, _sz],
    "fem_hip_qfunction_apply": [_vp, _i, _i, _vp, _vp, _i, _vp, _vp, _vp],
    "fem_hip_qfunction_destroy": [_vp],
    "fem_hip_operator_create": [_vp, ctypes.POINTER(_vp)],
    "fem_hip_operator_set_field": [_vp, ctypes.c_char_p, _vp, _vp, _i, _vp, _sz, _i],
    "fem_hip_operator_check_ready": [_vp],
    "fem_hip_operator_apply": [_vp, _vp, _sz, _vp, _sz, _vp],
    "fem_hip_operator_apply_f32": [_vp, _vp, _sz, _vp, _sz, _vp],
    "fem_hip_operator_apply_noalloc": [_vp, _vp, _sz, _vp, _sz, _vp],
    "fem_hip_noalloc_scratch_bytes": [],
    "fem_hip_operator_get_info": [_vp, _llp],
    "fem_hip_operator_set_fused_options": [_vp, _i, _i, _i],
    "fem_hip_operator_set_fused_algorithm": [_vp, _i, _i],
    "fem_hip_operator_set_box_layers": [_vp, _i],
    "fem_hip_operator_get_fused_choice": [_vp, _i, _ip, _ip],
    "fem_hip_operator_autotune": [_vp, _i, _i, _vp, _ip, _ip, ctypes.POINTER(_d), ctypes.POINTER(_d), _i, _ip],
    "fem_hip_tune_table_get": [_ip, ctypes.POINTER(_d), _llp, _i, _ip],
    "fem_hip_tune_table_set": [_ip, _d, ctypes.c_longlong],
    "fem_hip_tune_table_clear": [],
    "fem_hip_tune_table_save": [ctypes.c_char_p],
    "fem_hip_tune_table_load": [ctypes.c_char_p, _ip],
    "fem_hip_mesh_read": [ctypes.c_char_p, ctypes.c_void_p, _vp, _vp, _vp],
    "fem_hip_mesh_write": [ctypes.c_char_p, ctypes.c_void_p, _vp, _vp, _vp],
    "fem_hip_server_create": [ctypes.c_void_p, ctypes.POINTER(_vp)],
    "fem_hip_server_submit": [_vp, ctypes.c_void_p, _llp],
    "fem_hip_server_wait": [_vp, ctypes.c_longlong, _d, ctypes.c_void_p, _vp, _sz],
    "fem_hip_server_stats": [_vp, _llp],
    "fem_hip_server_destroy": [_vp],
    "fem_hip_serve_request_bytes": [ctypes.c_void_p, ctypes.POINTER(_sz)],
    "fem_hip_operator_mass_diagonal": [_vp, _vp, _sz, _vp],
    "fem_hip_operator_timing_begin": [_vp],
    "fem_hip_operator_timing_end": [_vp, ctypes.POINTER(_d), ctypes.POINTER(_d), _ip],
    "fem_hip_operator_destroy": [_vp],
    "fem_hip_operator_create_mass": [_vp, _vp, _vp, ctypes.POINTER(_vp)],
    "fem_hip_rng_uniform": [ctypes.c_uint64, _sz, _d, _d, _vp],
    "fem_hip_rng_uniform_device": [ctypes.c_uint64, _sz, _d, _d, _vp, _vp],
    "fem_hip_stream_copy": [_vp, _vp, _sz, _vp],
    "fem_hip_box_offsets": [_i, _i, _i, _i, _i, _vp],
    "fem_hip_box_coords": [_i, _d, ctypes.c_uint64, _vp],
    "fem_hip_box_offsets_device": [_i, _i, _i, _i, _i, _vp, _vp],
    "fem_hip_box_coords_device": [_i, _d, ctypes.c_uint64, _vp, _vp],
    "fem_hip_block_plan_stats": [_i, _i, _i, _vp, _i, _i, _i, _i, _llp],
    "fem_hip_box_detect": [_i, _i, _vp, _i, _i, _i, _vp, _vp],
    "fem_hip_box_column_check": [_i, _i, _i, _i, _i, _i, _llp],
    "fem_hip_mass_solve": [_vp, _vp, _vp, ctypes.POINTER(SolveOptions), ctypes.POINTER(SolveStats), _vp],
    "fem_hip_mass_solve_batched": [_i, ctypes.POINTER(_vp), ctypes.POINTER(_vp), ctypes.POINTER(_vp),
                                   ctypes.POINTER(SolveOptions), _i, ctypes.POINTER(SolveStats), _ip],
}
_RESTYPES = {"fem_hip_last_error": ctypes.c_char_p, "fem_hip_version": ctypes.c_char_p,
             "fem_hip_noalloc_scratch_bytes": ctypes.c_size_t}

_lib = None


def lib():
    """The loaded libfemhip; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise FemHipError(ERR_RUNTIME, "libfemhip.so is not built (" + LIB_PATH + "): run __graft_entry__.build()")
        L = ctypes.CDLL(LIB_PATH)
        for name, args in SIGNATURES.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = _RESTYPES.get(name, ctypes.c_int)
        _lib = L
    return _lib


def check(rc):
    if rc != SUCCESS:
        msg = lib().fem_hip_last_error().decode()
        if rc == ERR_INVALID_ARGUMENT:
            raise InvalidArgument(msg)
        raise FemHipError(rc, msg)


def call(name, *args):
    check(getattr(lib(), name)(*args))


def _ptr(a):
    return ctypes.c_void_p(a.ctypes.data) if a is not None else None


# ---------------------------------------------------------------- device


def device_count():
    n = ctypes.c_int(0)
    call("fem_hip_device_count", ctypes.byref(n))
    return n.value


def device_info(device=0):
    info = DeviceInfo()
    call("fem_hip_device_info_get", device, ctypes.byref(info))
    return info


def synchronize():
    call("fem_hip_device_synchronize")


class DeviceArray:
    """Owner of one device allocation (include/cuda/device-array.hpp:18-95)."""

    def __init__(self, n=0, dtype=np.float64):
        self.dtype = np.dtype(dtype)
        self.size = 0
        self.ptr = ctypes.c_void_p(None)
        self._alloc(int(n))

    @classmethod
    def from_host(cls, host, dtype=None):
        host = np.ascontiguousarray(host, dtype=dtype if dtype is not None else np.asarray(host).dtype)
        d = cls(host.size, host.dtype)
        d.copy_from_host(host)
        return d

    def _alloc(self, n):
        if n:
            p = ctypes.c_void_p()
            call("fem_hip_malloc", ctypes.byref(p), n * self.dtype.itemsize)
            self.ptr = p
        self.size = n

    def release(self):
        if self.ptr.value:
            lib().fem_hip_free(self.ptr)
        self.ptr = ctypes.c_void_p(None)
        self.size = 0

    def __del__(self):
        try:
            self.release()
        except Exception:
            pass

    @property
    def nbytes(self):
        return self.size * self.dtype.itemsize

    def resize(self, n):
        if n != self.size:
            self.release()
            self._alloc(int(n))

    def zero(self, stream=None):
        if self.size:
            call("fem_hip_memset_zero", self.ptr, self.nbytes, stream)

    def copy_from_host(self, host):
        host = np.ascontiguousarray(host, dtype=self.dtype)
        if host.size != self.size:
            raise InvalidArgument("DeviceArray::copy_from_host: size mismatch")
        if self.size:
            call("fem_hip_memcpy_htod", self.ptr, _ptr(host), self.nbytes)

    def to_host(self):
        out = np.empty(self.size, dtype=self.dtype)
        if self.size:
            call("fem_hip_memcpy_dtoh", _ptr(out), self.ptr, self.nbytes)
        return out


class Stream:
    def __init__(self):
        self.h = ctypes.c_void_p()
        call("fem_hip_stream_create", ctypes.byref(self.h))

    def synchronize(self):
        call("fem_hip_stream_synchronize", self.h)

    def __del__(self):
        try:
            if self.h.value:
                lib().fem_hip_stream_destroy(self.h)
        except Exception:
            pass


class Event:
    def __init__(self):
        self.h = ctypes.c_void_p()
        call("fem_hip_event_create", ctypes.byref(self.h))

    def record(self, stream=None):
        call("fem_hip_event_record", self.h, stream.h if isinstance(stream, Stream) else stream)

    def synchronize(self):
        call("fem_hip_event_synchronize", self.h)

    def elapsed_ms(self, later):
        ms = ctypes.c_float()
        call("fem_hip_event_elapsed_ms", self.h, later.h, ctypes.byref(ms))
        return ms.value

    def __del__(self):
        try:
            if self.h.value:
                lib().fem_hip_event_destroy(self.h)
        except Exception:
            pass


def _s(stream):
    return stream.h if isinstance(stream, Stream) else stream


# ---------------------------------------------------------------- host setup


def gauss_legendre(Q):
    x, w = np.zeros(Q), np.zeros(Q)
    call("fem_hip_gauss_legendre", Q, _ptr(x), _ptr(w))
    return x, w


def gauss_lobatto(Q, weights=True):
    x = np.zeros(Q)
    w = np.zeros(Q) if weights else None
    call("fem_hip_gauss_lobatto", Q, _ptr(x), _ptr(w))
    return (x, w) if weights else x


def lagrange_basis_matrix(nodes, pts):
    nodes = np.ascontiguousarray(nodes, np.float64)
    pts = np.ascontiguousarray(pts, np.float64)
    I = np.zeros(len(pts) * len(nodes))
    G = np.zeros_like(I)
    call("fem_hip_lagrange_basis_matrix", len(nodes), _ptr(nodes), len(pts), _ptr(pts), _ptr(I), _ptr(G))
    return I, G


def rng_uniform(seed, n, lo=-1.0, hi=1.0):
    out = np.empty(n)
    call("fem_hip_rng_uniform", seed, n, lo, hi, _ptr(out))
    return out


def stream_copy(dst, src, stream=None):
    """dst = src (device arrays of equal byte size, a multiple of 16) by the
    streaming copy kernel."""
    if dst.nbytes != src.nbytes:
        raise InvalidArgument("stream_copy: arrays differ in size")
    call("fem_hip_stream_copy", dst.ptr, src.ptr, src.nbytes, getattr(stream, "h", stream))


def rng_uniform_device(seed, n, lo=-1.0, hi=1.0, stream=None):
    """rng_uniform's stream, generated on the device into a new DeviceArray."""
    out = DeviceArray(n)
    call("fem_hip_rng_uniform_device", seed, n, lo, hi, out.ptr, _s(stream))
    return out


def box_offsets(nx, ny, nz, P, node_stride=1):
    out = np.empty(nx * ny * nz * P ** 3, dtype=np.int32)
    call("fem_hip_box_offsets", nx, ny, nz, P, node_stride, _ptr(out))
    return out


def box_offsets_device(nx, ny, nz, P, node_stride=1, stream=None):
    """box_offsets made on the device (a DeviceArray of int32)."""
    out = DeviceArray(nx * ny * nz * P ** 3, np.int32)
    call("fem_hip_box_offsets_device", nx, ny, nz, P, node_stride, out.ptr, _s(stream))
    return out


def box_coords_device(n, jitter=0.1, seed=0, stream=None):
    """box_coords made on the device (a DeviceArray of 3*(n+1)^3 doubles)."""
    out = DeviceArray(3 * (n + 1) ** 3)
    call("fem_hip_box_coords_device", n, float(jitter), ctypes.c_uint64(seed), out.ptr, _s(stream))
    return out


def box_coords(n, jitter=0.1, seed=0):
    out = np.empty(3 * (n + 1) ** 3)
    call("fem_hip_box_coords", n, jitter, seed, _ptr(out))
    return out


def block_plan_stats(num_elem, elem_size, l_size, offsets, nb, max_nodes=0, num_comp=1, comp_stride=None):
    offsets = np.ascontiguousarray(offsets, np.int32)
    st = (ctypes.c_longlong * 9)()
    cs = l_size if comp_stride is None else comp_stride
    call("fem_hip_block_plan_stats", num_elem, elem_size, l_size, _ptr(offsets), nb, max_nodes, num_comp, cs, st)
    keys = ["blocks", "max_nodes", "unique_nodes", "interior_nodes", "boundary_nodes", "element_nodes", "max_colors",
            "rows", "covers_all"]
    return dict(zip(keys, [int(x) for x in st]))


def box_detect(num_elem, P, offsets, num_comp=1, comp_stride=None, l_size=None):
    """(nx, ny, nz) if offsets are a lexicographic structured box of P^3-node hexes, else None."""
    offsets = np.ascontiguousarray(offsets, np.int32)
    n3 = (ctypes.c_int * 3)()
    ok = ctypes.c_int()
    n = int(offsets.max()) + 1
    cs = n if comp_stride is None else comp_stride
    ls = num_comp * n if l_size is None else l_size
    call("fem_hip_box_detect", num_elem, P, _ptr(offsets), num_comp, cs, ls, n3, ctypes.byref(ok))
    return tuple(n3) if ok.value else None


def box_column_check(nx, ny, nz, P, elems, layers=0):
    """Host check of the box-column index arithmetic (raises on a violation)."""
    st = (ctypes.c_longlong * 6)()
    call("fem_hip_box_column_check", nx, ny, nz, P, elems, layers, st)
    keys = ["segments", "layers", "shared_nodes", "partial_slots", "direct_nodes", "max_owners"]
    return dict(zip(keys, [int(x) for x in st]))


def tune_table():
    """Rows (P, Q, num_comp, fp32, deterministic, box, alg, elems, ms, num_elem) of the
    process-wide tuning table (box = 1: rows for structured boxes). A shape may
    hold rows measured on several mesh sizes (num_elem elements); an operator
    follows the row measured nearest its own size."""
    n = ctypes.c_int()
    call("fem_hip_tune_table_get", None, None, None, 0, ctypes.byref(n))
    k = n.value
    rows = (ctypes.c_int * (8 * max(k, 1)))()
    ms = (_d * max(k, 1))()
    ne = (ctypes.c_longlong * max(k, 1))()
    call("fem_hip_tune_table_get", rows, ms, ne, k, ctypes.byref(n))
    return [tuple(rows[8 * i:8 * i + 8]) + (float(ms[i]), int(ne[i])) for i in range(k)]


def tune_table_set(P, Q, num_comp, fp32, deterministic, alg, elems, ms, num_elem, box=0):
    """Adds the row for a shape measured on a mesh of num_elem elements (> 0),
    replacing the shape's row of that size."""
    row = (ctypes.c_int * 8)(P, Q, num_comp, int(fp32), int(deterministic), int(box), alg, elems)
    call("fem_hip_tune_table_set", row, float(ms), int(num_elem))


def tune_table_clear():
    call("fem_hip_tune_table_clear")


def tune_table_save(path):
    call("fem_hip_tune_table_save", os.fsencode(path))


def tune_table_load(path):
    n = ctypes.c_int()
    call("fem_hip_tune_table_load", os.fsencode(path), ctypes.byref(n))
    return n.value


TUNE_FILE = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "tuning", "mi355x_gfx950.txt")


def load_default_tuning():
    """Loads the committed MI355X tuning table if present; returns the rows taken."""
    return tune_table_load(TUNE_FILE) if os.path.exists(TUNE_FILE) else 0


def transpose_map(num_elem, elem_size, l_size, offsets):
    offsets = np.ascontiguousarray(offsets, np.int32)
    rows = ctypes.c_int()
    lvec = np.zeros(l_size, np.int32)
    toff = np.zeros(l_size + 1, np.int32)
    tind = np.zeros(num_elem * elem_size, np.int32)
    call("fem_hip_transpose_map_build", num_elem, elem_size, l_size, _ptr(offsets), ctypes.byref(rows),
         _ptr(lvec), _ptr(toff), _ptr(tind))
    n = rows.value
    return lvec[:n].copy(), toff[: n + 1].copy(), tind


# ---------------------------------------------------------------- restriction


class ElemRestriction:
    """Device element restriction (fem::cuda::ElemRestriction counterpart)."""

    def __init__(self, handle):
        self.h = handle
        s = (ctypes.c_longlong * 8)()
        call("fem_hip_restriction_get_sizes", self.h, s)
        (self.type, self.num_elem, self.elem_size, self.num_comp, self.comp_stride, self.l_size, self.e_size,
         self.num_nodes) = [int(x) for x in s]

    def __del__(self):
        try:
            if self.h and self.h.value:
                lib().fem_hip_restriction_destroy(self.h)
        except Exception:
            pass

    def transpose_map(self):
        lvec = np.zeros(self.num_nodes, np.int32)
        toff = np.zeros(self.num_nodes + 1, np.int32)
        tind = np.zeros(self.num_elem * self.elem_size, np.int32)
        call("fem_hip_restriction_get_transpose_map", self.h, _ptr(lvec), _ptr(toff), _ptr(tind))
        return lvec, toff, tind


def elem_restriction_create(num_elem, elem_size, num_comp, comp_stride, l_size, offsets,
                            scatter_mode=SCATTER_ATOMIC):
    offsets = np.ascontiguousarray(offsets, np.int32)
    if offsets.size != num_elem * elem_size:
        raise InvalidArgument("elem_restriction_create: offsets size must be num_elem * elem_size")
    h = ctypes.c_void_p()
    call("fem_hip_restriction_create", num_elem, elem_size, num_comp, comp_stride, l_size, _ptr(offsets),
         scatter_mode, ctypes.byref(h))
    return ElemRestriction(h)


def elem_restriction_create_strided(num_elem, elem_size, num_comp, l_size, strides=None):
    st = None if strides is None else np.ascontiguousarray(strides, np.int32)
    h = ctypes.c_void_p()
    call("fem_hip_restriction_create_strided", num_elem, elem_size, num_comp, l_size, _ptr(st), ctypes.byref(h))
    return ElemRestriction(h)


def elem_restriction_apply(r, t_mode, d_in, d_out, stream=None):
    name = "fem_hip_restriction_apply_f32" if d_in.dtype == np.float32 else "fem_hip_restriction_apply"
    call(name, r.h, t_mode, d_in.ptr, d_in.size, d_out.ptr, d_out.size, _s(stream))


# ---------------------------------------------------------------- basis


class TensorBasis:
    def __init__(self, handle):
        self.h = handle
        dims = (ctypes.c_int * 4)()
        call("fem_hip_basis_get", self.h, dims, None, None, None, None)
        self.dim, self.num_comp, self.P_1d, self.Q_1d = [int(x) for x in dims]
        P, Q = self.P_1d, self.Q_1d
        self.q_ref_1d, self.q_weight_1d = np.zeros(Q), np.zeros(Q)
        self.interp_1d, self.grad_1d = np.zeros(Q * P), np.zeros(Q * P)
        call("fem_hip_basis_get", self.h, None, _ptr(self.q_ref_1d), _ptr(self.q_weight_1d), _ptr(self.interp_1d),
             _ptr(self.grad_1d))

    def __del__(self):
        try:
            if self.h and self.h.value:
                lib().fem_hip_basis_destroy(self.h)
        except Exception:
            pass

    @classmethod
    def create_tensor_H1_lagrange(cls, dim, num_comp, P_1d, Q_1d, quad_mode=QUAD_GAUSS):
        h = ctypes.c_void_p()
        call("fem_hip_basis_create_h1_lagrange", dim, num_comp, P_1d, Q_1d, quad_mode, ctypes.byref(h))
        return cls(h)

    @classmethod
    def from_arrays(cls, dim, num_comp, P_1d, Q_1d, q_ref, q_weight, interp, grad):
        arrs = [np.ascontiguousarray(a, np.float64) for a in (q_ref, q_weight, interp, grad)]
        h = ctypes.c_void_p()
        call("fem_hip_basis_create_tensor", dim, num_comp, P_1d, Q_1d, *[_ptr(a) for a in arrs], ctypes.byref(h))
        return cls(h)


def tensor_basis_apply(b, num_elem, t_mode, eval_mode, d_u, d_v, stream=None):
    u_ptr = d_u.ptr if d_u is not None else None
    u_len = d_u.size if d_u is not None else 0
    call("fem_hip_basis_apply", b.h, num_elem, t_mode, eval_mode, u_ptr, u_len, d_v.ptr, d_v.size, _s(stream))


def tensor_contract_apply(A, B, C, J, d_t, t_mode, add, d_u, d_v, stream=None):
    call("fem_hip_contract", A, B, C, J, d_t.ptr, t_mode, int(add), d_u.ptr, d_v.ptr, _s(stream))


# ---------------------------------------------------------------- qfunction / operator


class QFunction:
    def __init__(self, kind):
        self.h = ctypes.c_void_p()
        call("fem_hip_qfunction_create", kind, ctypes.byref(self.h))
        self.inputs, self.outputs = [], []

    def __del__(self):
        try:
            if self.h.value:
                lib().fem_hip_qfunction_destroy(self.h)
        except Exception:
            pass

    def add_input(self, name, size, eval_mode):
        call("fem_hip_qfunction_add_input", self.h, name.encode(), size, eval_mode)
        self.inputs.append((name, size, eval_mode))

    def add_output(self, name, size, eval_mode):
        call("fem_hip_qfunction_add_output", self.h, name.encode(), size, eval_mode)
        self.outputs.append((name, size, eval_mode))

    def set_context_int(self, value):
        v = ctypes.c_int(value)
        call("fem_hip_qfunction_set_context", self.h, ctypes.byref(v), ctypes.sizeof(v))

    def set_context_ints(self, *values):
        v = (ctypes.c_int * len(values))(*values)
        call("fem_hip_qfunction_set_context", self.h, v, ctypes.sizeof(v))

    def apply(self, Q, d_in, d_out, stream=None):
        nin, nout = len(d_in), len(d_out)
        ip = (ctypes.c_void_p * max(nin, 1))(*[a.ptr.value for a in d_in])
        il = (ctypes.c_size_t * max(nin, 1))(*[a.size for a in d_in])
        op = (ctypes.c_void_p * max(nout, 1))(*[a.ptr.value for a in d_out])
        ol = (ctypes.c_size_t * max(nout, 1))(*[a.size for a in d_out])
        call("fem_hip_qfunction_apply", self.h, Q, nin, ip, il, nout, op, ol, _s(stream))


class Operator:
    def __init__(self, qf=None, handle=None):
        self.h = ctypes.c_void_p()
        if handle is not None:
            self.h = handle
        else:
            call("fem_hip_operator_create", qf.h, ctypes.byref(self.h))

    def __del__(self):
        try:
            if self.h.value:
                lib().fem_hip_operator_destroy(self.h)
        except Exception:
            pass

    def set_field(self, name, restriction, basis, vector_kind, passive=None):
        data, n, on_dev = None, 0, 0
        if isinstance(passive, DeviceArray):
            data, n, on_dev = passive.ptr, passive.size, 1
        elif passive is not None:
            self._keep = np.ascontiguousarray(passive, np.float64)
            data, n = _ptr(self._keep), self._keep.size
        call("fem_hip_operator_set_field", self.h, name.encode(), restriction.h if restriction else None,
             basis.h if basis else None, vector_kind, data, n, on_dev)

    def check_ready(self):
        call("fem_hip_operator_check_ready", self.h)

    def info(self):
        v = (ctypes.c_longlong * 8)()
        call("fem_hip_operator_get_info", self.h, v)
        keys = ["num_elem", "num_qpts", "active_in_l_size", "active_out_l_size", "fused", "P", "Q", "num_comp"]
        return dict(zip(keys, [int(x) for x in v]))

    def set_fused_options(self, scatter_mode=-1, exact=0, elems_per_block=0):
        call("fem_hip_operator_set_fused_options", self.h, scatter_mode, exact, elems_per_block)

    def set_fused_algorithm(self, alg, elems=0):
        call("fem_hip_operator_set_fused_algorithm", self.h, alg, elems)

    def set_box_layers(self, layers):
        """Element layers per workgroup of FUSED_BOX_COLUMN (0: automatic)."""
        call("fem_hip_operator_set_box_layers", self.h, layers)

    def fused_choice(self, fp32=False):
        a, e = ctypes.c_int(), ctypes.c_int()
        call("fem_hip_operator_get_fused_choice", self.h, int(fp32), ctypes.byref(a), ctypes.byref(e))
        return a.value, e.value

    def autotune(self, fp32=False, reps=10, stream=None):
        """Times every valid algorithm/packing; returns (best, [(alg, elems, ms), ...])."""
        a, e, ms, n = ctypes.c_int(), ctypes.c_int(), _d(), ctypes.c_int()
        cap = 256
        res = (_d * (3 * cap))()
        call("fem_hip_operator_autotune", self.h, int(fp32), reps, _s(stream), ctypes.byref(a), ctypes.byref(e),
             ctypes.byref(ms), res, cap, ctypes.byref(n))
        rows = [(int(res[3 * i]), int(res[3 * i + 1]), float(res[3 * i + 2])) for i in range(min(n.value, cap))]
        return (a.value, e.value, ms.value), rows

    def apply(self, d_in, d_out, stream=None):
        in_ptr = d_in.ptr if d_in is not None else None
        in_len = d_in.size if d_in is not None else 0
        if d_out.dtype == np.float32:
            call("fem_hip_operator_apply_f32", self.h, in_ptr, in_len, d_out.ptr, d_out.size, _s(stream))
        else:
            call("fem_hip_operator_apply", self.h, in_ptr, in_len, d_out.ptr, d_out.size, _s(stream))

    def apply_noalloc(self, d_in, d_out, stream=None):
        """FP64 apply that allocates no device memory (scratch in a library-owned
        device array; the operator's state must exist: apply it once first)."""
        call("fem_hip_operator_apply_noalloc", self.h, d_in.ptr, d_in.size, d_out.ptr, d_out.size, _s(stream))

    def timing_begin(self):
        call("fem_hip_operator_timing_begin", self.h)

    def timing_end(self):
        k, a, n = ctypes.c_double(), ctypes.c_double(), ctypes.c_int()
        call("fem_hip_operator_timing_end", self.h, ctypes.byref(k), ctypes.byref(a), ctypes.byref(n))
        return k.value, a.value, n.value

    def mass_diagonal(self, d_diag, stream=None):
        call("fem_hip_operator_mass_diagonal", self.h, d_diag.ptr, d_diag.size, _s(stream))

    def solve(self, d_b, d_x, mode=SOLVE_FP64, rtol=1e-10, max_iter=2000, max_outer=30, inner_rtol=None,
              stream=None, cg_variant=CG_AUTO, separate_second_pass=False):
        """M x = b. inner_rtol: reduction of each inner FP32 solve (SOLVE_MIXED, default 1e-5)
        or between residual replacements (SOLVE_MIXED_RELIABLE, default 1e-2). cg_variant:
        CG_AUTO, CG_PLAIN, CG_PIPELINED or CG_PLAIN_WIDE; separate_second_pass: the pipelined form keeps the
        apply's second pass as its own kernel instead of folding it into its step."""
        if inner_rtol is None:
            inner_rtol = 0.0  # the library's default for the mode
        opt = SolveOptions(mode, max_iter, max_outer, rtol, inner_rtol, cg_variant, int(separate_second_pass))
        st = SolveStats()
        call("fem_hip_mass_solve", self.h, d_b.ptr, d_x.ptr, ctypes.byref(opt), ctypes.byref(st), _s(stream))
        return {"iterations": st.iterations, "outer": st.outer, "rel_residual": st.rel_residual, "ms": st.ms}


def mass_solve_batched(ops, d_bs, d_xs, modes=SOLVE_FP64, rtol=1e-10, max_iter=2000, max_outer=30,
                       inner_rtol=None, iters_per_launch=8):
    """Solves M_i x_i = b_i for every operator together (fem_hip_mass_solve_batched):
    solves of one shape and precision share each persistent launch of their
    cohort's graph. modes: one mode or one per solve (SOLVE_FP64 / SOLVE_MIXED).
    Returns a list of stats dicts."""
    n = len(ops)
    if isinstance(modes, int):
        modes = [modes] * n
    opts = (SolveOptions * n)(*[SolveOptions(m, max_iter, max_outer, rtol, 0.0 if inner_rtol is None else inner_rtol,
                                             CG_AUTO, 0) for m in modes])
    sts = (SolveStats * n)()
    status = (ctypes.c_int * n)()
    hs = (_vp * n)(*[o.h for o in ops])
    bs = (_vp * n)(*[b.ptr for b in d_bs])
    xs = (_vp * n)(*[x.ptr for x in d_xs])
    call("fem_hip_mass_solve_batched", n, hs, bs, xs, opts, iters_per_launch, sts, status)
    return [{"iterations": st.iterations, "outer": st.outer, "rel_residual": st.rel_residual, "ms": st.ms}
            for st in sts]


def operator_create_mass(restriction, basis, d_qdata):
    h = ctypes.c_void_p()
    call("fem_hip_operator_create_mass", restriction.h, basis.h, d_qdata.ptr, ctypes.byref(h))
    return Operator(handle=h)


def operator_apply(op, d_in, d_out, stream=None):
    op.apply(d_in, d_out, stream)


# ---------------------------------------------------------------- on-disk meshes
class MeshInfo(ctypes.Structure):
    _fields_ = [("dim", ctypes.c_int), ("num_elem", ctypes.c_int), ("P", ctypes.c_int), ("num_nodes", ctypes.c_int),
                ("num_vertices", ctypes.c_int)]


def mesh_read(path):
    """A FEMHMSH1 mesh file as a dict (dim, num_elem, P, num_nodes, num_vertices,
    offsets [E*P^dim], vertex_offsets [E*2^dim], coords [dim*num_vertices])."""
    info = MeshInfo()
    call("fem_hip_mesh_read", os.fsencode(path), ctypes.byref(info), None, None, None)
    off = np.empty(info.num_elem * info.P ** info.dim, np.int32)
    voff = np.empty(info.num_elem * 2 ** info.dim, np.int32)
    coords = np.empty(info.dim * info.num_vertices)
    call("fem_hip_mesh_read", os.fsencode(path), ctypes.byref(info), _ptr(off), _ptr(voff), _ptr(coords))
    return {"dim": info.dim, "num_elem": info.num_elem, "P": info.P, "num_nodes": info.num_nodes,
            "num_vertices": info.num_vertices, "offsets": off, "vertex_offsets": voff, "coords": coords}


def mesh_write(path, mesh):
    info = MeshInfo(mesh["dim"], mesh["num_elem"], mesh["P"], mesh["num_nodes"], mesh["num_vertices"])
    off = np.ascontiguousarray(mesh["offsets"], np.int32)
    voff = np.ascontiguousarray(mesh["vertex_offsets"], np.int32)
    coords = np.ascontiguousarray(mesh["coords"], np.float64)
    call("fem_hip_mesh_write", os.fsencode(path), ctypes.byref(info), _ptr(off), _ptr(voff), _ptr(coords))


# ---------------------------------------------------------------- serving loop
class ServerOptions(ctypes.Structure):
    _fields_ = [("max_active", ctypes.c_int), ("workers", ctypes.c_int), ("hbm_fraction", ctypes.c_double),
                ("batch_iters", ctypes.c_int), ("autotune", ctypes.c_int), ("keep_x", ctypes.c_int),
                ("cheap_setup", ctypes.c_int), ("batched", ctypes.c_int), ("batch_max_elem_nodes", ctypes.c_longlong)]


class ServeRequest(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int), ("P", ctypes.c_int), ("Q", ctypes.c_int), ("num_comp", ctypes.c_int),
                ("mode", ctypes.c_int), ("jitter", ctypes.c_double), ("rtol", ctypes.c_double),
                ("mesh_seed", ctypes.c_uint64), ("rhs_seed", ctypes.c_uint64)]


class ServeResult(ctypes.Structure):
    _fields_ = [("status", ctypes.c_int), ("iterations", ctypes.c_int), ("outer", ctypes.c_int),
                ("rel_residual", ctypes.c_double), ("queue_ms", ctypes.c_double), ("setup_ms", ctypes.c_double),
                ("solve_ms", ctypes.c_double), ("latency_ms", ctypes.c_double), ("l_size", ctypes.c_longlong),
                ("est_bytes", ctypes.c_size_t), ("build_ms", ctypes.c_double), ("graph_ms", ctypes.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


def serve_request(n, P, Q=None, num_comp=1, mode=SOLVE_FP64, jitter=0.1, rtol=1e-10, mesh_seed=0, rhs_seed=1):
    return ServeRequest(n, P, P + 1 if Q is None else Q, num_comp, mode, jitter, rtol, mesh_seed, rhs_seed)


def serve_l_size(req):
    return req.num_comp * (req.n * (req.P - 1) + 1) ** 3


def serve_request_bytes(req):
    b = _sz()
    call("fem_hip_serve_request_bytes", ctypes.byref(req), ctypes.byref(b))
    return b.value


class Server:
    """Continuous-batching mass-solve server (fem_hip_server_*). batched: requests
    of one shape and precision with at most batch_max_elem_nodes element nodes
    (0: 2^20) share the persistent launches of their cohort's captured graph;
    others get one stream and graph per request."""

    def __init__(self, max_active=16, workers=8, hbm_fraction=0.5, batch_iters=8, autotune=False, keep_x=True,
                 cheap_setup=True, batched=True, batch_max_elem_nodes=0):
        o = ServerOptions(max_active, workers, hbm_fraction, batch_iters, int(autotune), int(keep_x),
                          int(cheap_setup), int(batched), int(batch_max_elem_nodes))
        self.h = ctypes.c_void_p()
        call("fem_hip_server_create", ctypes.byref(o), ctypes.byref(self.h))
        self.keep_x = keep_x

    def submit(self, req):
        i = ctypes.c_longlong()
        call("fem_hip_server_submit", self.h, ctypes.byref(req), ctypes.byref(i))
        return i.value

    def wait(self, rid, l_size=0, timeout_ms=-1.0):
        """(result dict, solution or None) once request `rid` is done, None on a
        timeout; raises if the request failed. l_size > 0 (server made with
        keep_x) also returns the solution of that length."""
        res = ServeResult()
        x = np.empty(l_size) if l_size else None
        rc = lib().fem_hip_server_wait(self.h, rid, float(timeout_ms), ctypes.byref(res),
                                       x.ctypes.data_as(_vp) if x is not None else None, l_size)
        if rc == ERR_NOT_READY:
            return None
        check(rc)
        return res.as_dict(), x

    def stats(self):
        st = (ctypes.c_longlong * 9)()
        call("fem_hip_server_stats", self.h, st)
        keys = ["submitted", "completed", "failed", "active", "queued", "active_peak", "budget", "used", "used_peak"]
        return dict(zip(keys, [int(v) for v in st]))

    def close(self):
        if self.h:
            call("fem_hip_server_destroy", self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
