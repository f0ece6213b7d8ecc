"""ctypes binding of the CPU oracle (oracle/liborc.so) -- test infrastructure.

Only tests/, ``__graft_entry__.smoke()`` and bench.py's ``cpu_baseline`` leg
import this module.  The solver product never does.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

FAR, NEIG, NARROW, NEXT, FROZEN, TAIL, BLOCK = range(7)
LOCAL_REF, LOCAL_EXACT = 0, 1
DMAX = np.finfo(np.float64).max
BIG = 1.0e300

_pd = C.POINTER(C.c_double)
_pi64 = C.POINTER(C.c_int64)
_pi32 = C.POINTER(C.c_int32)


class _Mesh(C.Structure):
    _fields_ = [("nn", C.c_int64), ("ne", C.c_int64), ("xyz", _pd),
                ("e2n", _pi64), ("fib", _pd), ("vel", _pd), ("atags", _pi32)]


class _Events(C.Structure):
    _fields_ = [("nev", C.c_int32), ("active", _pi32), ("tstart", _pd),
                ("dsp", _pi64), ("vtx", _pi64)]


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liborc.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle not built: run `make -C oracle`")
        L = C.CDLL(path)
        L.orc_tri_ref.restype = C.c_double
        L.orc_tri_ref.argtypes = [C.c_double] * 5
        L.orc_tet_ref.restype = C.c_double
        L.orc_tet_ref.argtypes = [C.c_double] * 9
        L.orc_tri_exact.restype = C.c_double
        L.orc_tri_exact.argtypes = [C.c_double] * 5
        L.orc_tet_exact.restype = C.c_double
        L.orc_tet_exact.argtypes = [C.c_double] * 9
        L.orc_local.restype = C.c_double
        L.orc_local.argtypes = [C.c_int, _pd, _pd, C.c_int]
        L.orc_relabel.restype = None
        L.orc_relabel.argtypes = [C.c_int, C.POINTER(C.c_int)]
        L.orc_metric.restype = None
        L.orc_metric.argtypes = [_pd, C.c_double, C.c_double, _pd]
        L.orc_geo_elem.restype = None
        L.orc_geo_elem.argtypes = [_pd, _pi64, _pd, _pd]
        L.orc_front_width.restype = C.c_double
        L.orc_front_width.argtypes = [C.c_int64, _pi64, _pd]
        L.orc_extract_numbering.restype = C.c_int64
        L.orc_extract_numbering.argtypes = [C.c_int64, _pi64, C.c_int64, _pi32,
                                            _pi64, _pi64, _pi64, _pi64]
        L.orc_connectivity.restype = C.c_int64
        L.orc_connectivity.argtypes = [C.c_int64, _pi64,
                                       C.POINTER(_pi32), C.POINTER(_pi64), C.POINTER(_pi64),
                                       C.POINTER(_pi32), C.POINTER(_pi64), C.POINTER(_pi64),
                                       _pi64]
        L.orc_free.restype = None
        L.orc_free.argtypes = [C.c_void_p]
        L.orc_compat_solve.restype = C.c_int
        L.orc_compat_solve.argtypes = [C.POINTER(_Mesh), C.POINTER(_Events), C.c_double,
                                       C.c_int64, C.c_double, C.c_double, C.c_int,
                                       _pd, _pi32, _pd]
        L.orc_compat_solve_par.restype = C.c_int
        L.orc_compat_solve_par.argtypes = [C.POINTER(_Mesh), C.POINTER(_Events), C.c_double,
                                           C.c_int64, C.c_double, C.c_double, C.c_int, C.c_int,
                                           _pd, _pi32, _pd]
        L.orc_compat_solve_timed.restype = C.c_int
        L.orc_compat_solve_timed.argtypes = [C.POINTER(_Mesh), C.POINTER(_Events), C.c_double,
                                       C.c_int64, C.c_double, C.c_double, C.c_int,
                                       _pd, _pi32, _pd]
        L.orc_fim_solve.restype = C.c_int
        L.orc_fim_solve.argtypes = [C.POINTER(_Mesh), C.c_int64, _pi64, _pd, C.c_int,
                                    _pd, _pd]
        L.orc_residual.restype = C.c_double
        L.orc_residual.argtypes = [C.POINTER(_Mesh), _pd, C.c_int]
        _LIB = L
    return _LIB


def _p(a, t):
    return a.ctypes.data_as(t) if a is not None else C.cast(None, t)


class Mesh:
    """Holds contiguous numpy arrays and the C view of them."""

    def __init__(self, xyz, e2n, fib=None, vel=None, atags=None):
        self.xyz = np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1, 3)
        self.e2n = np.ascontiguousarray(e2n, dtype=np.int64).reshape(-1, 4)
        self.fib = None if fib is None else np.ascontiguousarray(fib, dtype=np.float64).reshape(-1, 3)
        self.vel = None if vel is None else np.ascontiguousarray(vel, dtype=np.float64).reshape(-1, 2)
        self.atags = None if atags is None else np.ascontiguousarray(atags, dtype=np.int32)
        self.c = _Mesh(self.xyz.shape[0], self.e2n.shape[0], _p(self.xyz, _pd), _p(self.e2n, _pi64),
                       _p(self.fib, _pd), _p(self.vel, _pd), _p(self.atags, _pi32))


def tri_ref(e12, e13, e23, p1, p2):
    return lib().orc_tri_ref(e12, e13, e23, p1, p2)


def plant(which):
    """Test hook: plant mistake `which` in the reference tet formula (0 = none)."""
    lib().orc_plant(int(which))


def tet_ref(*a):
    return lib().orc_tet_ref(*a)


def tri_exact(e12, e14, e24, p1, p2):
    return lib().orc_tri_exact(e12, e14, e24, p1, p2)


def tet_exact(*a):
    return lib().orc_tet_exact(*a)


def local(variant, geo, ph, t):
    g = np.ascontiguousarray(geo, dtype=np.float64)
    p = np.ascontiguousarray(ph, dtype=np.float64)
    return lib().orc_local(variant, _p(g, _pd), _p(p, _pd), t)


def relabel(t):
    out = (C.c_int * 9)()
    lib().orc_relabel(t, out)
    return list(out)


def metric(f, vf, vs):
    f = np.ascontiguousarray(f, dtype=np.float64)
    M = np.zeros(6)
    lib().orc_metric(_p(f, _pd), vf, vs, _p(M, _pd))
    return M


def geo_elem(xyz, con, M):
    xyz = np.ascontiguousarray(xyz, dtype=np.float64)
    con = np.ascontiguousarray(con, dtype=np.int64)
    M = np.ascontiguousarray(M, dtype=np.float64)
    out = np.zeros(6)
    lib().orc_geo_elem(_p(xyz, _pd), _p(con, _pi64), _p(M, _pd), _p(out, _pd))
    return out


def front_width(e2n_sorted, xyz):
    e = np.ascontiguousarray(e2n_sorted, dtype=np.int64)
    x = np.ascontiguousarray(xyz, dtype=np.float64)
    return lib().orc_front_width(e.shape[0], _p(e, _pi64), _p(x, _pd))


def connectivity(e2n):
    """Returns sorted e2n and (n2e_cnt, n2e_dsp, n2e_con, n2n_cnt, n2n_dsp, n2n_con)."""
    e = np.array(e2n, dtype=np.int64).reshape(-1, 4).copy()
    ptrs = [_pi32(), _pi64(), _pi64(), _pi32(), _pi64(), _pi64()]
    tot = C.c_int64()
    L = lib()
    nn = L.orc_connectivity(e.shape[0], _p(e, _pi64), C.byref(ptrs[0]), C.byref(ptrs[1]),
                            C.byref(ptrs[2]), C.byref(ptrs[3]), C.byref(ptrs[4]),
                            C.byref(ptrs[5]), C.byref(tot))
    sizes = [nn, nn, 4 * e.shape[0], nn, nn, tot.value]
    out = []
    for p, n in zip(ptrs, sizes):
        out.append(np.ctypeslib.as_array(p, shape=(n,)).copy())
        L.orc_free(C.cast(p, C.c_void_p))
    return e, tuple(out)


def extract_numbering(e2n_sorted, nn, atags):
    e = np.ascontiguousarray(e2n_sorted, dtype=np.int64)
    t = np.ascontiguousarray(atags, dtype=np.int32)
    ne = e.shape[0]
    e2a = np.zeros(ne, dtype=np.int64)
    dsp = np.zeros(ne + 1, dtype=np.int64)
    tot = C.c_int64()
    L = lib()
    nagg = L.orc_extract_numbering(ne, _p(e, _pi64), nn, _p(t, _pi32), _p(e2a, _pi64),
                                   _p(dsp, _pi64), C.cast(None, _pi64), C.byref(tot))
    l2g = np.zeros(tot.value, dtype=np.int64)
    L.orc_extract_numbering(ne, _p(e, _pi64), nn, _p(t, _pi32), _p(e2a, _pi64),
                            _p(dsp, _pi64), _p(l2g, _pi64), C.byref(tot))
    return int(nagg), e2a, dsp[: nagg + 1].copy(), l2g


def _events(stims):
    """stims: list of (active, tstart_us, [vertices])"""
    act = np.array([s[0] for s in stims], dtype=np.int32)
    ts = np.array([s[1] for s in stims], dtype=np.float64)
    vt = [np.asarray(s[2], dtype=np.int64) for s in stims]
    dsp = np.zeros(len(stims) + 1, dtype=np.int64)
    dsp[1:] = np.cumsum([len(v) for v in vt]) if vt else []
    vtx = np.concatenate(vt) if vt else np.zeros(0, dtype=np.int64)
    keep = (act, ts, dsp, vtx)
    ev = _Events(len(stims), _p(act, _pi32), _p(ts, _pd), _p(dsp, _pi64), _p(vtx, _pi64))
    return ev, keep


def compat_solve(mesh: Mesh, stims, dt, n_steps, apd=3e5, front_width=-1.0,
                 variant=LOCAL_REF, timed=False):
    ev, keep = _events(stims)
    nn = mesh.xyz.shape[0]
    phi = np.zeros(nn)
    st = np.zeros(nn, dtype=np.int32)
    stats = np.zeros(2)
    fn = lib().orc_compat_solve_timed if timed else lib().orc_compat_solve
    rc = fn(C.byref(mesh.c), C.byref(ev), dt, n_steps, apd, front_width,
                                variant, _p(phi, _pd), _p(st, _pi32), _p(stats, _pd))
    if rc != 0:
        raise RuntimeError(f"orc_compat_solve failed: {rc}")
    del keep
    return phi, st, stats


def compat_solve_par(mesh: Mesh, stims, dt, n_steps, threads, apd=3e5, front_width=-1.0,
                     variant=LOCAL_REF):
    """compat_solve with the step loop on `threads` host threads; returns
    (phi, states, seconds) with seconds = (setup, step loop, merge)."""
    ev, keep = _events(stims)
    nn = mesh.xyz.shape[0]
    phi = np.zeros(nn)
    st = np.zeros(nn, dtype=np.int32)
    sec = np.zeros(3)
    rc = lib().orc_compat_solve_par(C.byref(mesh.c), C.byref(ev), dt, n_steps, apd, front_width,
                                    variant, threads, _p(phi, _pd), _p(st, _pi32), _p(sec, _pd))
    if rc != 0:
        raise RuntimeError(f"orc_compat_solve_par failed: {rc}")
    del keep
    return phi, st, sec


def fim_solve(mesh: Mesh, src, t0=None, variant=LOCAL_EXACT):
    src = np.ascontiguousarray(src, dtype=np.int64)
    t = None if t0 is None else np.ascontiguousarray(t0, dtype=np.float64)
    phi = np.zeros(mesh.xyz.shape[0])
    stats = np.zeros(2)
    rc = lib().orc_fim_solve(C.byref(mesh.c), src.shape[0], _p(src, _pi64), _p(t, _pd),
                             variant, _p(phi, _pd), _p(stats, _pd))
    if rc != 0:
        raise RuntimeError(f"orc_fim_solve failed: {rc}")
    return phi, stats


def residual(mesh: Mesh, phi, variant=LOCAL_EXACT):
    p = np.ascontiguousarray(phi, dtype=np.float64)
    return lib().orc_residual(C.byref(mesh.c), _p(p, _pd), variant)
