"""ctypes binding of the CPU oracle (oracle/liboracle.so, the restatement) and of
oracle/_ref/libfemref.so (the reference's own sources, built in the development
container only). Test infrastructure: only tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg load these.
"""
import ctypes
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "liboracle.so")
REF_SO = os.path.join(ROOT, "oracle", "_ref", "libfemref.so")

_vp, _i, _d, _sz = ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.c_size_t

_SIGS = {
    "orc_is_reference": ([], _i),
    "orc_last_error": ([], ctypes.c_char_p),
    "orc_gauss": ([_i, _vp, _vp], _i),
    "orc_lobatto": ([_i, _vp, _vp], _i),
    "orc_lagrange": ([_i, _vp, _i, _vp, _vp, _vp], _i),
    "orc_basis": ([_i, _i, _i, _i, _i, _vp, _vp, _vp, _vp], _i),
    "orc_contract": ([_i, _i, _i, _i, _vp, _i, _i, _i, _vp, _vp], _i),
    "orc_basis_apply": ([_i, _i, _i, _i, _i, _i, _i, _i, _vp, _sz, _vp, _sz, _i], _i),
    "orc_restriction_apply": ([_i, _i, _i, _i, _i, _vp, _vp, _i, _i, _vp, _vp], _i),
    "orc_restriction_multiplicity": ([_i, _i, _i, _i, _i, _vp, _vp], _i),
    "orc_qf_build_mass": ([_i, _i, _vp, _vp, _vp], _i),
    "orc_qf_apply_mass": ([_i, _vp, _vp, _vp], _i),
    "orc_build_qdata": ([_i, _i, _i, _i, _i, _i, _i, _vp, _vp, _vp], _i),
    "orc_mass_create": ([_i, _i, _i, _i, _i, _i, _i, _i, _vp, _vp], _vp),
    "orc_mass_apply": ([_vp, _vp, _vp], _i),
    "orc_mass_destroy": ([_vp], None),
    "orc_mass_diagonal": ([_vp, _vp], _i),
    "orc_mass_cg": ([_vp, _vp, _vp, _vp, _i, _d, ctypes.POINTER(_i), ctypes.POINTER(_d)], _i),
    "orc_transpose_map": ([_i, _i, _i, _vp, ctypes.POINTER(_i), _vp, _vp, _vp], _i),
    "orc_build_diffusion_qdata": ([_i, _i, _i, _i, _i, _i, _i, _vp, _vp, _vp], _i),
    "orc_diffusion_apply": ([_i, _i, _i, _i, _i, _i, _i, _i, _vp, _vp, _vp, _vp], _i),
}


class OracleError(Exception):
    def __init__(self, code, msg):
        super().__init__(msg)
        self.code = code


def _p(a):
    return ctypes.c_void_p(a.ctypes.data) if a is not None else None


class Oracle:
    """One loaded CPU library (restatement or reference)."""

    def __init__(self, path):
        if not os.path.exists(path):
            raise FileNotFoundError(path)
        self.L = ctypes.CDLL(path)
        for name, (args, res) in _SIGS.items():
            f = getattr(self.L, name, None)
            if f is None:
                continue
            f.argtypes = args
            f.restype = res
        self.is_reference = bool(self.L.orc_is_reference())

    def _chk(self, rc):
        if rc != 0:
            raise OracleError(rc, self.L.orc_last_error().decode())

    def gauss(self, Q):
        x, w = np.zeros(Q), np.zeros(Q)
        self._chk(self.L.orc_gauss(Q, _p(x), _p(w)))
        return x, w

    def lobatto(self, Q, weights=True):
        x = np.zeros(Q)
        w = np.zeros(Q) if weights else None
        self._chk(self.L.orc_lobatto(Q, _p(x), _p(w)))
        return x, w

    def lagrange(self, nodes, pts):
        nodes = np.ascontiguousarray(nodes, np.float64)
        pts = np.ascontiguousarray(pts, np.float64)
        I, G = np.zeros(len(nodes) * len(pts)), np.zeros(len(nodes) * len(pts))
        self._chk(self.L.orc_lagrange(len(nodes), _p(nodes), len(pts), _p(pts), _p(I), _p(G)))
        return I, G

    def basis(self, dim, nc, P, Q, quad_mode=0):
        a = [np.zeros(Q), np.zeros(Q), np.zeros(Q * P), np.zeros(Q * P)]
        self._chk(self.L.orc_basis(dim, nc, P, Q, quad_mode, *[_p(x) for x in a]))
        return dict(q_ref=a[0], q_weight=a[1], interp=a[2], grad=a[3])

    def contract(self, A, B, C, J, t, t_mode, add, u, v=None):
        t = np.ascontiguousarray(t, np.float64)
        u = np.ascontiguousarray(u, np.float64)
        out = np.zeros(A * J * C) if v is None else np.array(v, np.float64)
        self._chk(self.L.orc_contract(A, B, C, J, _p(t), t.size, t_mode, int(add), _p(u), _p(out)))
        return out

    def basis_apply(self, dim, nc, P, Q, num_elem, t_mode, eval_mode, u, out_len, quad_mode=0, v_in=None):
        u = np.ascontiguousarray(u if u is not None else np.zeros(0), np.float64)
        v = np.zeros(out_len) if v_in is None else np.array(v_in, np.float64)
        self._chk(self.L.orc_basis_apply(dim, nc, P, Q, quad_mode, num_elem, t_mode, eval_mode, _p(u), u.size,
                                         _p(v), out_len, 0 if v_in is None else 1))
        return v

    def restriction_apply(self, ne, es, nc, cs, l_size, offsets, t_mode, inp, out=None, strides=None,
                          backend=False):
        off = None if offsets is None else np.ascontiguousarray(offsets, np.int32)
        st = None if strides is None else np.ascontiguousarray(strides, np.int32)
        inp = np.ascontiguousarray(inp, np.float64)
        if t_mode == 0:
            o = np.zeros(nc * es * ne)
        else:
            o = np.zeros(l_size) if out is None else np.array(out, np.float64)
        self._chk(self.L.orc_restriction_apply(ne, es, nc, cs, l_size, _p(off), _p(st), int(backend), t_mode,
                                               _p(inp), _p(o)))
        return o

    def multiplicity(self, ne, es, nc, cs, l_size, offsets):
        off = np.ascontiguousarray(offsets, np.int32)
        m = np.zeros(l_size)
        self._chk(self.L.orc_restriction_multiplicity(ne, es, nc, cs, l_size, _p(off), _p(m)))
        return m

    def qf_build_mass(self, dim, Q, J, w):
        J = np.ascontiguousarray(J, np.float64)
        w = np.ascontiguousarray(w, np.float64)
        out = np.zeros(Q)
        self._chk(self.L.orc_qf_build_mass(dim, Q, _p(J), _p(w), _p(out)))
        return out

    def qf_apply_mass(self, Q, u, qd):
        u = np.ascontiguousarray(u, np.float64)
        qd = np.ascontiguousarray(qd, np.float64)
        out = np.zeros(Q)
        self._chk(self.L.orc_qf_apply_mass(Q, _p(u), _p(qd), _p(out)))
        return out

    def build_qdata(self, dim, ne, mesh_P, Q, coord_l_size, coord_comp_stride, mesh_offsets, coords, quad_mode=0):
        mo = np.ascontiguousarray(mesh_offsets, np.int32)
        co = np.ascontiguousarray(coords, np.float64)
        out = np.zeros(ne * Q ** dim)
        self._chk(self.L.orc_build_qdata(dim, ne, mesh_P, Q, quad_mode, coord_l_size, coord_comp_stride, _p(mo),
                                         _p(co), _p(out)))
        return out

    def build_diffusion_qdata(self, dim, ne, mesh_P, Q, coord_l_size, coord_comp_stride, mesh_offsets, coords,
                              quad_mode=0):
        mo = np.ascontiguousarray(mesh_offsets, np.int32)
        co = np.ascontiguousarray(coords, np.float64)
        out = np.zeros(dim * (dim + 1) // 2 * ne * Q ** dim)
        self._chk(self.L.orc_build_diffusion_qdata(dim, ne, mesh_P, Q, quad_mode, coord_l_size, coord_comp_stride,
                                                   _p(mo), _p(co), _p(out)))
        return out

    def diffusion_apply(self, dim, ne, P, Q, nc, l_size, comp_stride, offsets, qdata, u, quad_mode=0):
        off = np.ascontiguousarray(offsets, np.int32)
        qd = np.ascontiguousarray(qdata, np.float64)
        uu = np.ascontiguousarray(u, np.float64)
        out = np.zeros(l_size)
        self._chk(self.L.orc_diffusion_apply(dim, ne, P, Q, nc, quad_mode, l_size, comp_stride, _p(off), _p(qd),
                                             _p(uu), _p(out)))
        return out

    def mass(self, dim, ne, P, Q, nc, l_size, comp_stride, offsets, qdata, quad_mode=0):
        return MassOp(self, dim, ne, P, Q, nc, l_size, comp_stride, offsets, qdata, quad_mode)

    def transpose_map(self, ne, es, l_size, offsets):
        off = np.ascontiguousarray(offsets, np.int32)
        rows = ctypes.c_int()
        lvec = np.zeros(l_size, np.int32)
        toff = np.zeros(l_size + 1, np.int32)
        tind = np.zeros(ne * es, np.int32)
        self._chk(self.L.orc_transpose_map(ne, es, l_size, _p(off), ctypes.byref(rows), _p(lvec), _p(toff),
                                           _p(tind)))
        n = rows.value
        return lvec[:n].copy(), toff[: n + 1].copy(), tind


class MassOp:
    def __init__(self, orc, dim, ne, P, Q, nc, l_size, comp_stride, offsets, qdata, quad_mode=0):
        self.orc = orc
        self.l_size = l_size
        off = np.ascontiguousarray(offsets, np.int32)
        qd = np.ascontiguousarray(qdata, np.float64)
        self.h = orc.L.orc_mass_create(dim, ne, P, Q, nc, quad_mode, l_size, comp_stride, _p(off), _p(qd))
        if not self.h:
            raise OracleError(1, orc.L.orc_last_error().decode())

    def __del__(self):
        try:
            self.orc.L.orc_mass_destroy(self.h)
        except Exception:
            pass

    def apply(self, u):
        u = np.ascontiguousarray(u, np.float64)
        v = np.zeros(self.l_size)
        self.orc._chk(self.orc.L.orc_mass_apply(self.h, _p(u), _p(v)))
        return v

    def diagonal(self):
        d = np.zeros(self.l_size)
        self.orc._chk(self.orc.L.orc_mass_diagonal(self.h, _p(d)))
        return d

    def cg(self, b, diag=None, max_iter=2000, rtol=1e-10):
        b = np.ascontiguousarray(b, np.float64)
        x = np.zeros(self.l_size)
        it, rel = ctypes.c_int(), ctypes.c_double()
        dg = None if diag is None else np.ascontiguousarray(diag, np.float64)
        self.orc._chk(self.orc.L.orc_mass_cg(self.h, _p(b), _p(dg), _p(x), max_iter, rtol, ctypes.byref(it),
                                             ctypes.byref(rel)))
        return x, it.value, rel.value


_cache = {}


def oracle():
    if "o" not in _cache:
        _cache["o"] = Oracle(ORACLE_SO)
    return _cache["o"]


def reference():
    """The reference's own CPU library, or None where it is not built."""
    if "r" not in _cache:
        _cache["r"] = Oracle(REF_SO) if os.path.exists(REF_SO) else None
    return _cache["r"]
