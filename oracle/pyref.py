"""ctypes binding of oracle/_ref/libref_host.so -- the reference's own host
code (connectivity, submesh extraction, geometry, VTK, config/plan parsing)
compiled from its sources by `make -C oracle ref`.  Test infrastructure only.

``available()`` is False where the library was not built (it needs the
reference sources at build time); tests that use it skip then.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_PATH = os.path.join(_HERE, "_ref", "libref_host.so")
_LIB = None

_pd = C.POINTER(C.c_double)
_pi64 = C.POINTER(C.c_int64)
_pi32 = C.POINTER(C.c_int32)


def available() -> bool:
    return os.path.exists(_PATH)


def lib():
    global _LIB
    if _LIB is None:
        L = C.CDLL(_PATH)
        L.ref_load_vtk.argtypes = [C.c_char_p, C.c_char_p, _pi64, _pi64]
        L.ref_mesh_get.argtypes = [_pd, _pi64, _pi32]
        L.ref_mesh_set.argtypes = [C.c_int64, C.c_int64, _pd, _pi64, _pi32]
        L.ref_connectivity.restype = C.c_int64
        L.ref_conn_get.argtypes = [_pi64, _pi32, _pi64, _pi64, _pi32, _pi64, _pi64]
        L.ref_geo.argtypes = [_pd]
        L.ref_extract.restype = C.c_int64
        L.ref_sub_sizes.argtypes = [C.c_int64, _pi64, _pi64, _pi64]
        L.ref_sub_get.argtypes = [C.c_int64, _pi64, _pi64, _pd, _pi32, _pi64, _pi32, _pi64]
        L.ref_n2i_size.restype = C.c_int64
        L.ref_n2i_get.argtypes = [_pi64, _pi64, _pi64]
        L.ref_save_vtk.argtypes = [C.c_char_p, _pd, _pd]
        L.ref_load_options.argtypes = [C.c_char_p, _pi32, _pi32, _pi32, _pd, _pi64]
        L.ref_load_plan.argtypes = [C.c_char_p, _pd, _pi64, _pi64]
        L.ref_plan_events.argtypes = [_pi32, _pd, _pd, _pi64, _pi64]
        _LIB = L
    return _LIB


def _p(a, t):
    return a.ctypes.data_as(t) if a is not None else C.cast(None, t)


def load_vtk(path, tags=None):
    nn, ne = C.c_int64(), C.c_int64()
    rc = lib().ref_load_vtk(path.encode(), (tags or "").encode(), C.byref(nn), C.byref(ne))
    if rc != 0:
        raise RuntimeError("ref_load_vtk failed")
    xyz = np.zeros((nn.value, 3))
    e2n = np.zeros((ne.value, 4), dtype=np.int64)
    tag = np.zeros(ne.value, dtype=np.int32)
    lib().ref_mesh_get(_p(xyz, _pd), _p(e2n, _pi64), _p(tag, _pi32))
    return xyz, e2n, tag


def set_mesh(xyz, e2n, atags=None):
    xyz = np.ascontiguousarray(xyz, dtype=np.float64)
    e2n = np.ascontiguousarray(e2n, dtype=np.int64)
    at = None if atags is None else np.ascontiguousarray(atags, dtype=np.int32)
    lib().ref_mesh_set(xyz.shape[0], e2n.shape[0], _p(xyz, _pd), _p(e2n, _pi64), _p(at, _pi32))


def connectivity(nn, ne):
    """Runs compute_base_connectivity on the current mesh."""
    tot = lib().ref_connectivity()
    e2n = np.zeros((ne, 4), dtype=np.int64)
    a = [np.zeros(nn, np.int32), np.zeros(nn, np.int64), np.zeros(4 * ne, np.int64),
         np.zeros(nn, np.int32), np.zeros(nn, np.int64), np.zeros(tot, np.int64)]
    lib().ref_conn_get(_p(e2n, _pi64), _p(a[0], _pi32), _p(a[1], _pi64), _p(a[2], _pi64),
                       _p(a[3], _pi32), _p(a[4], _pi64), _p(a[5], _pi64))
    return e2n, tuple(a)


def geo(ne):
    g = np.zeros(6 * ne)
    lib().ref_geo(_p(g, _pd))
    return g.reshape(ne, 6)


def extract(nn_global):
    """Runs extract_submeshes2 (after connectivity). Returns list of dicts and
    n2i as (count per global vertex, local ids, submesh ids)."""
    nagg = lib().ref_extract()
    if nagg < 0:
        raise RuntimeError("extract_submeshes2 failed")
    subs = []
    for a in range(nagg):
        nn, ne, nnnc = C.c_int64(), C.c_int64(), C.c_int64()
        lib().ref_sub_sizes(a, C.byref(nn), C.byref(ne), C.byref(nnnc))
        d = dict(l2g=np.zeros(nn.value, np.int64), e2n=np.zeros(4 * ne.value, np.int64),
                 geo=np.zeros(6 * ne.value), n2e_cnt=np.zeros(nn.value, np.int32),
                 n2e_con=np.zeros(4 * ne.value, np.int64), n2n_cnt=np.zeros(nn.value, np.int32),
                 n2n_con=np.zeros(nnnc.value, np.int64))
        lib().ref_sub_get(a, _p(d["l2g"], _pi64), _p(d["e2n"], _pi64), _p(d["geo"], _pd),
                          _p(d["n2e_cnt"], _pi32), _p(d["n2e_con"], _pi64),
                          _p(d["n2n_cnt"], _pi32), _p(d["n2n_con"], _pi64))
        subs.append(d)
    tot = lib().ref_n2i_size()
    cnt = np.zeros(nn_global, np.int64)
    ldx = np.zeros(tot, np.int64)
    adx = np.zeros(tot, np.int64)
    lib().ref_n2i_get(_p(cnt, _pi64), _p(ldx, _pi64), _p(adx, _pi64))
    return subs, (cnt, ldx, adx)


def save_vtk(path, phi=None, states=None):
    p = None if phi is None else np.ascontiguousarray(phi, dtype=np.float64)
    s = None if states is None else np.ascontiguousarray(states, dtype=np.float64)
    if lib().ref_save_vtk(path.encode(), _p(p, _pd), _p(s, _pd)) != 0:
        raise RuntimeError("ref_save_vtk failed")


def load_options(cfg):
    v, st, re = C.c_int32(), C.c_int32(), C.c_int32()
    dt, ns = C.c_double(), C.c_int64()
    rc = lib().ref_load_options(cfg.encode(), C.byref(v), C.byref(st), C.byref(re),
                                C.byref(dt), C.byref(ns))
    if rc != 0:
        raise RuntimeError("load_options failed")
    return dict(verbose=v.value, solvetype=st.value, rent_en=re.value, dt=dt.value,
                n_steps=ns.value)


def load_plan(plan):
    dt, ns, tv = C.c_double(), C.c_int64(), C.c_int64()
    nev = lib().ref_load_plan(plan.encode(), C.byref(dt), C.byref(ns), C.byref(tv))
    if nev < 0:
        raise RuntimeError("load_plan failed")
    act = np.zeros(nev, np.int32)
    ts = np.zeros(nev)
    du = np.zeros(nev)
    cnt = np.zeros(nev, np.int64)
    vtx = np.zeros(tv.value, np.int64)
    lib().ref_plan_events(_p(act, _pi32), _p(ts, _pd), _p(du, _pd), _p(cnt, _pi64), _p(vtx, _pi64))
    ev = []
    k = 0
    for i in range(nev):
        ev.append(dict(active=int(act[i]), tstart=float(ts[i]), dur=float(du[i]),
                       vtx=vtx[k:k + cnt[i]].copy()))
        k += cnt[i]
    return dict(dt=dt.value, n_steps=ns.value, events=ev)
