"""fimpy -- ctypes binding of libcufim.so (include/cufim.h).

The solver runs only through the HIP library: loading fails loudly when
``cufim-nvlink_amd/lib/libcufim.so`` has not been built, and no call falls
back to CPU code.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.dirname(_HERE)
# CUFIM_LIB selects another build of the library (same ABI)
LIB_PATH = os.environ.get("CUFIM_LIB") or os.path.join(PKG, "lib", "libcufim.so")
CLI_PATH = os.path.join(PKG, "bin", "cufim")

FAR, NEIG, NARROW, NEXT, FROZEN, TAIL = range(6)
MODE_CONVERGE, MODE_COMPAT = 0, 1
LOCAL_EXACT, LOCAL_REF = 0, 1
PART_TAGS, PART_RCB = 0, 1
DMAX = np.finfo(np.float64).max

ERRORS = {-1: "FIM_EINVAL", -2: "FIM_EIO", -3: "FIM_ELIMIT", -4: "FIM_EDEVICE",
          -5: "FIM_ENOCONV", -6: "FIM_ESTATE"}

_pd = C.POINTER(C.c_double)
_pf = C.POINTER(C.c_float)
_pi64 = C.POINTER(C.c_int64)
_pi32 = C.POINTER(C.c_int32)
_pu8 = C.POINTER(C.c_uint8)
_vp = C.c_void_p


class FimError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


class PartOpts(C.Structure):
    _fields_ = [("source", C.c_int32), ("max_nn", C.c_int32), ("max_ne", C.c_int32),
                ("seed", C.c_uint64)]


class PackInfo(C.Structure):
    _fields_ = [("nn_global", C.c_int64), ("ne_global", C.c_int64), ("nsub", C.c_int64),
                ("ncopy", C.c_int64), ("n_iface_copies", C.c_int64), ("nn_max", C.c_int32),
                ("ne_max", C.c_int32), ("device_bytes_mesh", C.c_int64), ("lds_bytes", C.c_int64)]


class Opts(C.Structure):
    _fields_ = [("mode", C.c_int32), ("local", C.c_int32), ("n_instances", C.c_int32),
                ("device", C.c_int32), ("block", C.c_int32), ("local_iters", C.c_int32),
                ("graph_iters", C.c_int32), ("grid", C.c_int32), ("eps_rel", C.c_double),
                ("apd", C.c_double), ("profile", C.c_int32), ("staged", C.c_int32),
                ("window", C.c_double), ("unit", C.c_int32), ("waves", C.c_int32)]


class NNOpts(C.Structure):
    _fields_ = [("device", C.c_int32), ("seed", C.c_uint64), ("lr", C.c_double),
                ("beta1", C.c_double), ("beta2", C.c_double), ("adam_eps", C.c_double),
                ("t_scale", C.c_double), ("max_maps", C.c_int32), ("max_batch", C.c_int32)]


NN_INPUTS, NN_HIDDEN, NN_LAYERS, NN_PARAMS = 16, 256, 6, 333569


class CfgOptions(C.Structure):
    _fields_ = [("verbose", C.c_int32), ("method_id", C.c_int32), ("solvetype", C.c_int32),
                ("rent_en", C.c_int32), ("dt", C.c_double), ("n_steps", C.c_int64)]


class Stats(C.Structure):
    _fields_ = [("global_iters", C.c_int64), ("activations", C.c_int64),
                ("local_iters", C.c_int64), ("elem_updates", C.c_int64), ("last_ms", C.c_double),
                ("alg_bytes", C.c_int64), ("compute_launches", C.c_int64), ("compute_ms", C.c_double),
                ("vertex_updates", C.c_int64)]


# every symbol include/cufim.h declares, with its ctypes signature
SIGNATURES = {
    "fim_last_error": (C.c_char_p, []),
    "fim_mesh_create": (C.c_int, [C.c_int64, C.c_int64, _pd, _pi64, _pd, _pd, _pi32, C.POINTER(_vp)]),
    "fim_mesh_load_vtk": (C.c_int, [C.c_char_p, C.POINTER(_vp)]),
    "fim_mesh_read_tags": (C.c_int, [_vp, C.c_char_p]),
    "fim_mesh_set_tags": (C.c_int, [_vp, _pi32]),
    "fim_mesh_set_regions": (C.c_int, [_vp, _pu8]),
    "fim_mesh_info": (C.c_int, [_vp, _pi64, _pi64]),
    "fim_mesh_get": (C.c_int, [_vp, _pd, _pi64, _pi32]),
    "fim_mesh_get_material": (C.c_int, [_vp, _pd, _pd]),
    "fim_mesh_front_width": (C.c_double, [_vp]),
    "fim_mesh_save_vtk": (C.c_int, [_vp, C.c_char_p, _pd, _pi32]),
    "fim_mesh_destroy": (None, [_vp]),
    "fim_mesh_gen_box": (C.c_int, [C.c_double, C.c_double, C.c_double, C.c_int, C.c_int, C.c_int,
                                   C.c_double, C.c_uint64, C.c_double, C.c_double, C.c_double,
                                   C.c_double, C.POINTER(_vp)]),
    "fim_mesh_concat": (C.c_int, [C.POINTER(_vp), C.c_int32, C.POINTER(_vp)]),
    "fim_mesh_gen_shell": (C.c_int, [C.c_double, C.c_double, C.c_double, C.c_double, C.c_double,
                                     C.c_int32, C.c_double, C.c_double, C.c_uint64, C.c_double,
                                     C.c_double, C.c_double, C.POINTER(_vp)]),
    "fim_pack_build": (C.c_int, [_vp, C.POINTER(PartOpts), C.POINTER(_vp)]),
    "fim_pack_info": (C.c_int, [_vp, C.POINTER(PackInfo)]),
    "fim_pack_partition": (C.c_int, [_vp, _pi64, _pi64]),
    "fim_pack_destroy": (None, [_vp]),
    "fim_pack_save_tags": (C.c_int, [_vp, C.c_char_p]),
    "fim_mesh_apply_plan_speeds": (C.c_int, [_vp, C.c_char_p, _pi64]),
    "fim_create": (C.c_int, [_vp, C.POINTER(Opts), C.POINTER(_vp)]),
    "fim_reset": (C.c_int, [_vp]),
    "fim_seed": (C.c_int, [_vp, C.c_int32, _pi64, C.c_int64, C.c_double]),
    "fim_seed_batch": (C.c_int, [_vp, C.c_int32, _pi32, _pi64, _pi64, C.c_double]),
    "fim_advance": (C.c_int, [_vp, C.c_int32, C.c_int32, _pi32]),
    "fim_run_multi": (C.c_int, [C.POINTER(_vp), C.c_int32, C.c_int32, _pi32]),
    "fim_active": (C.c_int, [_vp, _pi32]),
    "fim_recycle": (C.c_int, [_vp, C.c_int32, _pi32]),
    "fim_set_scaling": (C.c_int, [_vp, C.c_int32, _pf]),
    "fim_init_field": (C.c_int, [_vp, C.c_int32, _pf]),
    "fim_set_window": (C.c_int, [_vp, C.c_int32, C.c_double]),
    "fim_step": (C.c_int, [_vp, C.c_double, C.c_double, C.c_double, C.c_double, _pi32]),
    "fim_run": (C.c_int, [_vp, C.c_int32, _pi32]),
    "fim_download": (C.c_int, [_vp, C.c_int32, _pd, _pi32]),
    "fim_residual": (C.c_int, [_vp, C.c_int32, _pd]),
    "fim_support": (C.c_int, [_vp, C.c_int32, _pd]),
    "fim_support_noalloc": (C.c_int, [_vp, C.c_int32, _pd]),
    "fim_nn_create": (C.c_int, [_vp, C.POINTER(NNOpts), C.POINTER(_vp)]),
    "fim_nn_destroy": (None, [_vp]),
    "fim_nn_get_params": (C.c_int, [_vp, _pf]),
    "fim_nn_set_params": (C.c_int, [_vp, _pf]),
    "fim_nn_adam_step": (C.c_int, [_vp, _pf]),
    "fim_nn_get_moments": (C.c_int, [_vp, _pf, _pf, _pi64]),
    "fim_nn_features": (C.c_int, [_vp, _pi64, _pi64, C.c_int32, _pf, C.c_int64, _pf]),
    "fim_nn_add_maps": (C.c_int, [_vp, _vp, C.c_int32, C.c_int32, _pi64, _pf]),
    "fim_nn_dataset_size": (C.c_int, [_vp, _pi32]),
    "fim_nn_gradient": (C.c_int, [_vp, _pf, _pf, C.c_int32, _pd, _pf]),
    "fim_nn_train": (C.c_int, [_vp, C.c_int32, C.c_int32, C.c_uint64, _pd]),
    "fim_nn_predict": (C.c_int, [_vp, _pf, C.c_int64, _pf]),
    "fim_nn_predict_map": (C.c_int, [_vp, _pi64, C.c_int32, _pf, _pf]),
    "fim_nn_predict_max": (C.c_int, [_vp, C.c_int32, _pi64, _pf, _pi64, C.c_int32, _pf]),
    "fim_nn_warm_start": (C.c_int, [_vp, _vp, C.c_int32, _pi64, C.c_int32, _pf, C.c_double, C.c_int32]),
    "fim_stats": (C.c_int, [_vp, C.POINTER(Stats)]),
    "fim_destroy": (None, [_vp]),
    "fim_device_info": (C.c_int, [C.c_int32, C.c_char_p, C.c_int64]),
    "fim_config_load_options": (C.c_int, [C.c_char_p, C.POINTER(CfgOptions)]),
    "fim_config_load_plan": (C.c_int, [C.c_char_p, C.POINTER(CfgOptions), _pi32, _pi64]),
    "fim_config_plan_events": (C.c_int, [C.c_char_p, _pi32, _pd, _pd, _pi64, _pi64]),
    "fim_run_reference": (C.c_int, [C.c_char_p, C.c_char_p, C.c_char_p, C.c_char_p, C.c_int32,
                                    C.c_char_p, C.c_int32, _pd]),
}

_LIB = None


def lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} not built: run `python cufim-nvlink_amd/build.py`")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _LIB = L
    return _LIB


def _check(rc):
    if rc != 0:
        raise FimError(rc, lib().fim_last_error().decode())


def _p(a, t):
    return a.ctypes.data_as(t) if a is not None else C.cast(None, t)


class Mesh:
    def __init__(self, handle):
        self.h = handle
        nn, ne = C.c_int64(), C.c_int64()
        _check(lib().fim_mesh_info(self.h, C.byref(nn), C.byref(ne)))
        self.nn, self.ne = nn.value, ne.value

    @classmethod
    def from_arrays(cls, xyz, e2n, fib=None, vel=None, etags=None):
        xyz = np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1, 3)
        e2n = np.ascontiguousarray(e2n, dtype=np.int64).reshape(-1, 4)
        fib = None if fib is None else np.ascontiguousarray(fib, dtype=np.float64)
        vel = None if vel is None else np.ascontiguousarray(vel, dtype=np.float64)
        etags = None if etags is None else np.ascontiguousarray(etags, dtype=np.int32)
        h = _vp()
        _check(lib().fim_mesh_create(xyz.shape[0], e2n.shape[0], _p(xyz, _pd), _p(e2n, _pi64),
                                     _p(fib, _pd), _p(vel, _pd), _p(etags, _pi32), C.byref(h)))
        return cls(h)

    @classmethod
    def load_vtk(cls, path):
        h = _vp()
        _check(lib().fim_mesh_load_vtk(path.encode(), C.byref(h)))
        return cls(h)

    @classmethod
    def gen_box(cls, lx, ly, lz, nx, ny, nz, jitter=0.1, seed=0, angle0=0.0, angle1=0.0,
                vf=0.6, vc=0.6):
        h = _vp()
        _check(lib().fim_mesh_gen_box(lx, ly, lz, nx, ny, nz, jitter, seed, angle0, angle1, vf, vc,
                                      C.byref(h)))
        return cls(h)

    @classmethod
    def gen_shell(cls, ao, bo, ai, bi, h, n_scars=12, r0=3000.0, r1=10000.0, seed=0, vf=0.6, vc=0.24,
                  vs=0.05):
        hh = _vp()
        _check(lib().fim_mesh_gen_shell(ao, bo, ai, bi, h, n_scars, r0, r1, seed, vf, vc, vs, C.byref(hh)))
        return cls(hh)

    @classmethod
    def concat(cls, parts):
        arr = (_vp * len(parts))(*[q.h for q in parts])
        h = _vp()
        _check(lib().fim_mesh_concat(arr, len(parts), C.byref(h)))
        return cls(h)

    def read_tags(self, path):
        _check(lib().fim_mesh_read_tags(self.h, path.encode()))

    def set_tags(self, tags):
        t = np.ascontiguousarray(tags, dtype=np.int32)
        assert t.shape[0] == self.ne
        _check(lib().fim_mesh_set_tags(self.h, _p(t, _pi32)))

    def set_regions(self, reg):
        r = np.ascontiguousarray(reg, dtype=np.uint8)
        assert r.shape[0] == self.ne
        _check(lib().fim_mesh_set_regions(self.h, _p(r, _pu8)))

    def arrays(self):
        xyz = np.zeros((self.nn, 3))
        e2n = np.zeros((self.ne, 4), dtype=np.int64)
        tags = np.zeros(self.ne, dtype=np.int32)
        _check(lib().fim_mesh_get(self.h, _p(xyz, _pd), _p(e2n, _pi64), _p(tags, _pi32)))
        return xyz, e2n, tags

    def material(self):
        fib = np.zeros((self.ne, 3))
        vel = np.zeros((self.ne, 2))
        _check(lib().fim_mesh_get_material(self.h, _p(fib, _pd), _p(vel, _pd)))
        return fib, vel

    def front_width(self):
        return lib().fim_mesh_front_width(self.h)

    def apply_plan_speeds(self, plan_path):
        n = C.c_int64()
        _check(lib().fim_mesh_apply_plan_speeds(self.h, plan_path.encode(), C.byref(n)))
        return n.value

    def save_vtk(self, path, phi=None, state=None):
        p = None if phi is None else np.ascontiguousarray(phi, dtype=np.float64)
        s = None if state is None else np.ascontiguousarray(state, dtype=np.int32)
        _check(lib().fim_mesh_save_vtk(self.h, path.encode(), _p(p, _pd), _p(s, _pi32)))

    def __del__(self):
        if getattr(self, "h", None) and _LIB is not None:
            _LIB.fim_mesh_destroy(self.h)
            self.h = None


class Pack:
    def __init__(self, mesh: Mesh, source=PART_RCB, max_nn=512, max_ne=0, seed=0):
        o = PartOpts(source, max_nn, max_ne, seed)
        self.h = _vp()
        _check(lib().fim_pack_build(mesh.h, C.byref(o), C.byref(self.h)))
        i = PackInfo()
        _check(lib().fim_pack_info(self.h, C.byref(i)))
        self.info = {k: getattr(i, k) for k, _ in PackInfo._fields_}

    def partition(self):
        e2s = np.zeros(self.info["ne_global"], dtype=np.int64)
        c2g = np.zeros(self.info["ncopy"], dtype=np.int64)
        _check(lib().fim_pack_partition(self.h, _p(e2s, _pi64), _p(c2g, _pi64)))
        return e2s, c2g

    def save_tags(self, path):
        _check(lib().fim_pack_save_tags(self.h, path.encode()))

    def __del__(self):
        if getattr(self, "h", None) and _LIB is not None:
            _LIB.fim_pack_destroy(self.h)
            self.h = None


class Solver:
    def __init__(self, pack: Pack, mode=MODE_CONVERGE, local=LOCAL_EXACT, n_instances=1, device=0,
                 block=0, local_iters=0, graph_iters=0, grid=0, eps_rel=0.0, apd=0.0,
                 profile=False, window=0.0, staged=False, unit=0, waves=0):
        self.pack = pack
        self.nn = pack.info["nn_global"]
        self.n_instances = n_instances
        self.window = window  # every instance's window until set_window changes it
        o = Opts(mode, local, n_instances, device, block, local_iters, graph_iters, grid, eps_rel, apd,
                 int(profile), int(staged), window, unit, waves)
        self.h = _vp()
        _check(lib().fim_create(pack.h, C.byref(o), C.byref(self.h)))

    def reset(self):
        _check(lib().fim_reset(self.h))

    def seed(self, vtx, t0=0.0, inst=0):
        v = np.ascontiguousarray(vtx, dtype=np.int64).reshape(-1)
        _check(lib().fim_seed(self.h, inst, _p(v, _pi64), v.shape[0], t0))

    def seed_batch(self, sets, t0=0.0, insts=None):
        """Seeds instance insts[i] (default i) with the vertices sets[i], all
        in one transfer and one launch."""
        ins = np.arange(len(sets), dtype=np.int32) if insts is None else \
            np.ascontiguousarray(insts, dtype=np.int32)
        assert ins.shape[0] == len(sets)
        off = np.zeros(len(sets) + 1, dtype=np.int64)
        off[1:] = np.cumsum([len(np.asarray(v).reshape(-1)) for v in sets])
        v = np.ascontiguousarray(np.concatenate([np.asarray(x, dtype=np.int64).reshape(-1) for x in sets])
                                 if len(sets) else np.zeros(0, np.int64))
        _check(lib().fim_seed_batch(self.h, len(sets), _p(ins, _pi32), _p(off, _pi64), _p(v, _pi64), t0))

    def set_scaling(self, scale, inst=0):
        s = np.ascontiguousarray(scale, dtype=np.float32)
        assert s.shape == (4,)
        _check(lib().fim_set_scaling(self.h, inst, _p(s, _pf)))

    def set_window(self, window, inst=0):
        """Time window (us per global iteration) of one instance; 0 = off."""
        _check(lib().fim_set_window(self.h, inst, float(window)))

    def init_field(self, phi0, inst=0):
        p = np.ascontiguousarray(phi0, dtype=np.float32)
        assert p.shape[0] == self.nn
        _check(lib().fim_init_field(self.h, inst, _p(p, _pf)))

    def step(self, phi_start, cur_time, dt, front_width):
        n = C.c_int32()
        _check(lib().fim_step(self.h, phi_start, cur_time, dt, front_width, C.byref(n)))
        return n.value

    def run(self, max_iters=100000):
        it = C.c_int32()
        _check(lib().fim_run(self.h, max_iters, C.byref(it)))
        return it.value

    def advance(self, iters, reset_stats=False):
        """Runs at least `iters` global iterations (or until nothing is
        queued); returns the queued work units left."""
        left = C.c_int32()
        _check(lib().fim_advance(self.h, iters, int(reset_stats), C.byref(left)))
        return left.value

    def active(self):
        """Per instance: True while it still has queued work."""
        a = np.zeros(self.n_instances, dtype=np.int32)
        _check(lib().fim_active(self.h, _p(a, _pi32)))
        return a.astype(bool)

    def recycle(self, insts):
        """Clears converged instances so they can be seeded again while the
        others keep running."""
        v = np.ascontiguousarray(insts, dtype=np.int32).reshape(-1)
        _check(lib().fim_recycle(self.h, v.shape[0], _p(v, _pi32)))

    def download(self, inst=0):
        phi = np.zeros(self.nn)
        st = np.zeros(self.nn, dtype=np.int32)
        _check(lib().fim_download(self.h, inst, _p(phi, _pd), _p(st, _pi32)))
        return phi, st

    def residual(self, inst=0):
        r = C.c_double()
        _check(lib().fim_residual(self.h, inst, C.byref(r)))
        return r.value

    def support(self, inst=0):
        r = C.c_double()
        _check(lib().fim_support(self.h, inst, C.byref(r)))
        return r.value

    def support_noalloc(self, inst=0):
        """support() through the entry point that allocates nothing"""
        r = C.c_double()
        _check(lib().fim_support_noalloc(self.h, inst, C.byref(r)))
        return r.value

    def stats(self):
        s = Stats()
        _check(lib().fim_stats(self.h, C.byref(s)))
        return {k: getattr(s, k) for k, _ in Stats._fields_}

    def close(self):
        if getattr(self, "h", None) and _LIB is not None:
            _LIB.fim_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()


def device_info(device=0):
    buf = C.create_string_buffer(512)
    _check(lib().fim_device_info(device, buf, 512))
    return buf.value.decode()


def load_options(path):
    o = CfgOptions()
    _check(lib().fim_config_load_options(path.encode(), C.byref(o)))
    return {k: getattr(o, k) for k, _ in CfgOptions._fields_}


def load_plan(path, opts=None):
    """Returns (options with the plan's dt / n_steps, events)."""
    o = CfgOptions(**opts) if opts else CfgOptions()
    ne, nv = C.c_int32(), C.c_int64()
    _check(lib().fim_config_load_plan(path.encode(), C.byref(o), C.byref(ne), C.byref(nv)))
    act = np.zeros(ne.value, np.int32)
    ts = np.zeros(ne.value)
    du = np.zeros(ne.value)
    cnt = np.zeros(ne.value, np.int64)
    vtx = np.zeros(nv.value, np.int64)
    _check(lib().fim_config_plan_events(path.encode(), _p(act, _pi32), _p(ts, _pd), _p(du, _pd),
                                        _p(cnt, _pi64), _p(vtx, _pi64)))
    events, k = [], 0
    for i in range(ne.value):
        events.append(dict(active=int(act[i]), tstart=float(ts[i]), dur=float(du[i]),
                           vtx=vtx[k:k + cnt[i]].copy()))
        k += cnt[i]
    return {k: getattr(o, k) for k, _ in CfgOptions._fields_}, events


def run_many(solvers, max_iters=100000):
    """fim_run for several converge-mode contexts at once (their streams
    overlap on the device); returns each one's global iterations."""
    n = len(solvers)
    hs = (_vp * n)(*[s.h for s in solvers])
    it = (C.c_int32 * n)()
    _check(lib().fim_run_multi(hs, n, max_iters, it))
    return list(it)


class SolverGroup:
    """Converge-mode contexts solved together (fim_run_multi): instance i of
    the group is instance i - off[k] of context k.  Splitting a batch over two
    contexts lets one context's work fill the other's launch tails; the maps
    are the same as from one context."""

    def __init__(self, solvers):
        self.solvers = list(solvers)
        assert self.solvers
        self.off = np.cumsum([0] + [s.n_instances for s in self.solvers]).astype(np.int64)
        self.window = self.solvers[0].window

    @property
    def n_instances(self):
        return int(self.off[-1])

    def parts(self, n):
        """(context, first, end) of the group's instances 0..n-1, per context"""
        return [(s, int(self.off[k]), int(min(self.off[k + 1], n)))
                for k, s in enumerate(self.solvers) if self.off[k] < n]

    def _at(self, inst):
        if not 0 <= inst < self.n_instances:
            raise FimError(-1, f"instance {inst} out of range")
        k = int(np.searchsorted(self.off, inst, side="right")) - 1
        return self.solvers[k], int(inst - self.off[k])

    def download(self, inst=0):
        s, j = self._at(inst)
        return s.download(inst=j)

    def residual(self, inst=0):
        s, j = self._at(inst)
        return s.residual(inst=j)

    def run(self, max_iters=100000):
        return max(run_many(self.solvers, max_iters))

    def stats(self):
        """counters summed over the contexts; global iterations and device
        time of the longest"""
        sts = [s.stats() for s in self.solvers]
        out = {k: sum(st[k] for st in sts) for k in sts[0]}
        out["global_iters"] = max(st["global_iters"] for st in sts)
        out["last_ms"] = max(st["last_ms"] for st in sts)
        return out

    def close(self):
        for s in self.solvers:
            s.close()


def run_reference(mesh, part, plan, out, reps=1, cfg="data/default.cfg", device=0):
    s = C.c_double()
    _check(lib().fim_run_reference(mesh.encode(), part.encode(), plan.encode(),
                                   out.encode() if out else b"", reps, cfg.encode(), device,
                                   C.byref(s)))
    return s.value


def _sites3(sites):
    s = np.full(3, -1, dtype=np.int64)
    v = np.asarray(sites, dtype=np.int64).reshape(-1)
    assert 1 <= v.shape[0] <= 3
    s[:v.shape[0]] = v
    return s, int(v.shape[0])


class Surrogate:
    """Coordinate MLP surrogate (include/cufim.h, fim_nn_*)."""

    def __init__(self, mesh: Mesh, device=0, seed=0, lr=0.0, t_scale=0.0, max_maps=0, max_batch=0):
        self.mesh = mesh
        self.nn = mesh.nn
        o = NNOpts(device, seed, lr, 0.0, 0.0, 0.0, t_scale, max_maps, max_batch)
        self.h = _vp()
        _check(lib().fim_nn_create(mesh.h, C.byref(o), C.byref(self.h)))

    def params(self):
        p = np.zeros(NN_PARAMS, dtype=np.float32)
        _check(lib().fim_nn_get_params(self.h, _p(p, _pf)))
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        assert p.shape == (NN_PARAMS,)
        _check(lib().fim_nn_set_params(self.h, _p(p, _pf)))

    def adam_step(self, grad):
        """One Adam step of the masters with a caller-given gradient."""
        g = np.ascontiguousarray(grad, dtype=np.float32)
        assert g.shape == (NN_PARAMS,)
        _check(lib().fim_nn_adam_step(self.h, _p(g, _pf)))

    def moments(self):
        """(m, v, steps) of the optimiser, parameter order."""
        m = np.zeros(NN_PARAMS, dtype=np.float32)
        v = np.zeros(NN_PARAMS, dtype=np.float32)
        t = C.c_int64()
        _check(lib().fim_nn_get_moments(self.h, _p(m, _pf), _p(v, _pf), C.byref(t)))
        return m, v, t.value

    def features(self, vtx, sites, scales):
        v = np.ascontiguousarray(vtx, dtype=np.int64).reshape(-1)
        s, ns = _sites3(sites)
        sc = np.ascontiguousarray(scales, dtype=np.float32)
        x = np.zeros((v.shape[0], NN_INPUTS), dtype=np.float32)
        _check(lib().fim_nn_features(self.h, _p(v, _pi64), _p(s, _pi64), ns, _p(sc, _pf), v.shape[0],
                                     _p(x, _pf)))
        return x

    def add_maps(self, solver, inst0, sites, scales):
        """sites: list of 1-3 vertex ids per map; scales: (n, 4)"""
        n = len(sites)
        s = np.stack([_sites3(x)[0] for x in sites]).astype(np.int64)
        sc = np.ascontiguousarray(scales, dtype=np.float32).reshape(n, 4)
        _check(lib().fim_nn_add_maps(self.h, solver.h, inst0, n, _p(s, _pi64), _p(sc, _pf)))

    def dataset_size(self):
        n = C.c_int32()
        _check(lib().fim_nn_dataset_size(self.h, C.byref(n)))
        return n.value

    def gradient(self, x, t):
        x = np.ascontiguousarray(x, dtype=np.float32)
        t = np.ascontiguousarray(t, dtype=np.float32)
        g = np.zeros(NN_PARAMS, dtype=np.float32)
        loss = C.c_double()
        _check(lib().fim_nn_gradient(self.h, _p(x, _pf), _p(t, _pf), x.shape[0], C.byref(loss), _p(g, _pf)))
        return loss.value, g

    def train(self, steps, batch=65536, seed=0):
        loss = C.c_double()
        _check(lib().fim_nn_train(self.h, steps, batch, seed, C.byref(loss)))
        return loss.value

    def predict(self, x):
        x = np.ascontiguousarray(x, dtype=np.float32)
        t = np.zeros(x.shape[0], dtype=np.float32)
        _check(lib().fim_nn_predict(self.h, _p(x, _pf), x.shape[0], _p(t, _pf)))
        return t

    def predict_map(self, sites, scales):
        s, ns = _sites3(sites)
        sc = np.ascontiguousarray(scales, dtype=np.float32)
        t = np.zeros(self.nn, dtype=np.float32)
        _check(lib().fim_nn_predict_map(self.h, _p(s, _pi64), ns, _p(sc, _pf), _p(t, _pf)))
        return t

    def predict_max(self, solves, sample):
        """Latest predicted activation time of each (sites, scales) solve over
        the vertex sample."""
        n = len(solves)
        s = np.stack([_sites3(x[0])[0] for x in solves]).astype(np.int64) if n else np.zeros((0, 3), np.int64)
        sc = np.ascontiguousarray(np.stack([np.asarray(x[1], np.float32) for x in solves]) if n
                                  else np.zeros((0, 4), np.float32))
        v = np.ascontiguousarray(sample, dtype=np.int64).reshape(-1)
        out = np.zeros(n, dtype=np.float32)
        _check(lib().fim_nn_predict_max(self.h, n, _p(s, _pi64), _p(sc, _pf), _p(v, _pi64), v.shape[0],
                                        _p(out, _pf)))
        return out

    def warm_start(self, solver, inst, sites, scales, delta, all_active=True):
        s, ns = _sites3(sites)
        sc = np.ascontiguousarray(scales, dtype=np.float32)
        _check(lib().fim_nn_warm_start(self.h, solver.h, inst, _p(s, _pi64), ns, _p(sc, _pf), delta,
                                       int(all_active)))

    def close(self):
        if getattr(self, "h", None) and _LIB is not None:
            _LIB.fim_nn_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()
