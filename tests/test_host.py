"""Host side of the product, CPU only: the C ABI exports, mesh / VTK / config
parsing against the reference's own host code (oracle/_ref), the partitioner
and error paths.  No GPU call is made here."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

import cases
import fimpy as F
import pyoracle as O
import pyref as R

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
need_ref = pytest.mark.skipif(not R.available(), reason="oracle/_ref not built (needs reference sources)")
UTV = f"{cases.UT}/ut_mesh.vtk"
UTT = f"{cases.UT}/ut_mesh.tags"


def test_library_exports_every_declared_symbol():
    text = open(os.path.join(ROOT, "include", "cufim.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = set(re.findall(r"\b(fim_[a-z_0-9]+)\s*\(", text))
    assert len(names) > 25
    lib = ctypes.CDLL(F.LIB_PATH)
    for n in sorted(names):
        assert hasattr(lib, n), n
    assert names == set(F.SIGNATURES), names ^ set(F.SIGNATURES)


def test_option_structs_match_the_header(tmp_path):
    # the ctypes mirrors of the option structs have the C layout
    src = tmp_path / "sz.c"
    src.write_text('#include <stddef.h>\n#include <stdio.h>\n#include "cufim.h"\n'
                   'int main(void){printf("%zu %zu %zu %zu %zu\\n", sizeof(fim_opts),'
                   ' offsetof(fim_opts, window), offsetof(fim_opts, waves),'
                   ' sizeof(fim_part_opts), sizeof(fim_nn_opts));return 0;}\n')
    exe = tmp_path / "sz"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    got = [int(x) for x in subprocess.run([str(exe)], capture_output=True, text=True,
                                          check=True).stdout.split()]
    assert got == [ctypes.sizeof(F.Opts), F.Opts.window.offset, F.Opts.waves.offset,
                   ctypes.sizeof(F.PartOpts), ctypes.sizeof(F.NNOpts)]


@need_ref
def test_load_vtk_matches_reference():
    m = F.Mesh.load_vtk(UTV)
    xyz, e2n, tags = m.arrays()
    rx, re2n, rtag = R.load_vtk(UTV)
    R.connectivity(rx.shape[0], re2n.shape[0])
    assert np.array_equal(xyz, rx)
    rs, _ = R.connectivity(rx.shape[0], re2n.shape[0])
    assert np.array_equal(e2n, rs)
    assert np.array_equal(tags, rtag)


@need_ref
def test_save_vtk_bytes_match_reference(tmp_path):
    m = F.Mesh.load_vtk(UTV)
    rng = np.random.default_rng(0)
    phi = rng.uniform(0, 5e5, m.nn)
    phi[::7] = np.finfo(np.float64).max  # unreached -> -1
    st = rng.integers(0, 6, m.nn).astype(np.int32)
    ours = tmp_path / "ours.vtk"
    theirs = tmp_path / "ref.vtk"
    m.save_vtk(str(ours), phi, st)
    R.load_vtk(UTV)
    xyz, cells, _ = R.load_vtk(UTV)
    R.connectivity(xyz.shape[0], cells.shape[0])
    R.save_vtk(str(theirs), phi, st.astype(np.float64))
    assert ours.read_bytes() == theirs.read_bytes()


def test_save_vtk_format_of_golden():
    """the reference's output file is reproduced byte for byte from its own
    phi / states: the writer alone is pinned by the golden"""
    import vtkio
    g = vtkio.read_vtk(f"{cases.UT}/cuda_fim_res.vtk")
    m = F.Mesh.load_vtk(UTV)
    phi = np.where(g["point_data"]["phi"] < 0, np.finfo(np.float64).max, g["point_data"]["phi"])
    st = g["point_data"]["states"].astype(np.int32)
    import tempfile
    with tempfile.TemporaryDirectory() as d:
        p = os.path.join(d, "o.vtk")
        m.save_vtk(p, phi, st)
        assert open(p).read() == open(f"{cases.UT}/cuda_fim_res.vtk").read()


@need_ref
def test_tag_partition_numbering_matches_reference():
    for tagfile in ("ut_mesh.tags", "ut_mesh_p.tags"):
        m = F.Mesh.load_vtk(UTV)
        m.read_tags(f"{cases.UT}/{tagfile}")
        p = F.Pack(m, F.PART_TAGS)
        e2s, c2g = p.partition()
        xyz, cells, _ = R.load_vtk(UTV)
        R.set_mesh(xyz, cells, np.loadtxt(f"{cases.UT}/{tagfile}", dtype=np.int32))
        R.connectivity(xyz.shape[0], cells.shape[0])
        subs, _ = R.extract(xyz.shape[0])
        assert np.array_equal(c2g, np.concatenate([s["l2g"] for s in subs]))


@need_ref
def test_options_and_plan_match_reference():
    cfg = os.path.join(ROOT, "data", "default.cfg")
    ours = F.load_options(cfg)
    theirs = R.load_options(cfg)
    for k in ("verbose", "solvetype", "rent_en", "dt", "n_steps"):
        assert ours[k] == theirs[k], k
    o2, ev = F.load_plan(f"{cases.UT}/ut_mesh.plan.json", ours)
    rp = R.load_plan(f"{cases.UT}/ut_mesh.plan.json")
    assert o2["dt"] == rp["dt"] and o2["n_steps"] == rp["n_steps"]
    assert len(ev) == len(rp["events"])
    for a, b in zip(ev, rp["events"]):
        assert a["active"] == b["active"] and a["tstart"] == b["tstart"] and a["dur"] == b["dur"]
        assert np.array_equal(a["vtx"], b["vtx"])


def test_plan_pulse_trains_and_errors(tmp_path):
    import json
    plan = json.load(open(f"{cases.UT}/ut_mesh.plan.json"))
    st = plan["simulation"]["stimuli"]["S0"]
    st["protocol"].update(npls=3, bcl=400.0, start=5.0)
    plan["simulation"]["stimuli"]["A_off"] = dict(st, active=0)
    p = tmp_path / "p.json"
    p.write_text(json.dumps(plan))
    o, ev = F.load_plan(str(p))
    assert o["dt"] == 150e3 and o["n_steps"] == 4
    assert [e["tstart"] for e in ev] == [5e3, 405e3, 805e3]
    plan["simulation"]["stimuli"]["S0"]["type"] = "Extracellular"
    p.write_text(json.dumps(plan))
    with pytest.raises(F.FimError, match="unknown stim type"):
        F.load_plan(str(p))
    del plan["simulation"]["stimuli"]
    p.write_text(json.dumps(plan))
    assert F.load_plan(str(p))[1] == []
    with pytest.raises(F.FimError, match="FIM_EIO"):
        F.load_plan(str(tmp_path / "missing.json"))
    (tmp_path / "bad.cfg").write_text("true\nfmm\nvol\nconst\nmaybe\nfalse\nfrozen\n1.0\n10\n")
    with pytest.raises(F.FimError, match="re_flag"):
        F.load_options(str(tmp_path / "bad.cfg"))


def test_rcb_partition_is_deterministic_and_valid():
    m = F.Mesh.gen_box(seed=3, **cases.scaled(cases.SLAB, 4))
    a = F.Pack(m, F.PART_RCB, max_nn=300, seed=11)
    b = F.Pack(m, F.PART_RCB, max_nn=300, seed=11)
    ea, ca = a.partition()
    eb, cb = b.partition()
    assert np.array_equal(ea, eb) and np.array_equal(ca, cb)
    info = a.info
    assert info["nn_max"] <= 300
    assert np.bincount(ea).min() > 0 and ea.max() + 1 == info["nsub"]
    # every vertex has a copy in each submesh holding one of its elements
    _, e2n, _ = m.arrays()
    for s in range(0, info["nsub"], 7):
        verts = np.unique(e2n[ea == s])
        assert len(verts) <= 300
    c = F.Pack(m, F.PART_RCB, max_nn=300, seed=12)
    assert c.info["nsub"] > 0


def test_box_generator_is_seeded_and_conforming():
    cfg = dict(cases.CUBE, nx=5, ny=4, nz=3)
    a = F.Mesh.gen_box(seed=1, **cfg)
    b = F.Mesh.gen_box(seed=1, **cfg)
    c = F.Mesh.gen_box(seed=2, **cfg)
    xa, ea, _ = a.arrays()
    xb, eb, _ = b.arrays()
    xc, _, _ = c.arrays()
    assert np.array_equal(xa, xb) and np.array_equal(ea, eb) and not np.array_equal(xa, xc)
    assert a.ne == 6 * 5 * 4 * 3 and a.nn == 6 * 5 * 4
    # positive volumes
    v = np.einsum("ij,ij->i", xa[ea[:, 1]] - xa[ea[:, 0]],
                  np.cross(xa[ea[:, 2]] - xa[ea[:, 0]], xa[ea[:, 3]] - xa[ea[:, 0]]))
    assert np.all(np.abs(v) > 0)
    # conforming: every face shared by at most two elements, boundary faces
    # exactly on the box surface
    faces = {}
    for t in ea:
        for k in range(4):
            f = tuple(sorted(np.delete(t, k)))
            faces[f] = faces.get(f, 0) + 1
    assert max(faces.values()) == 2
    nb = sum(1 for v in faces.values() if v == 1)
    assert nb == 2 * 2 * (5 * 4 + 4 * 3 + 5 * 3)
    # jitter bound
    h = np.array([cfg["lx"] / 5, cfg["ly"] / 4, cfg["lz"] / 3])
    grid = np.stack(np.meshgrid(np.arange(6), np.arange(5), np.arange(4), indexing="ij"), -1)
    grid = grid.transpose(2, 1, 0, 3).reshape(-1, 3) * h
    assert np.all(np.abs(xa - grid) <= 0.1 * h + 1e-9)


def test_geometry_matches_oracle_metric():
    """product metric geometry (via the pack) agrees with the oracle's
    closed form e^T M e, checked through orc_geo_elem on the same input"""
    m = F.Mesh.gen_box(seed=5, **dict(cases.SLAB, nx=4, ny=4, nz=2))
    xyz, e2n, _ = m.arrays()
    fib, vel = m.material()
    for e in range(0, e2n.shape[0], 5):
        M = O.metric(fib[e], vel[e, 0], vel[e, 1])
        g = O.geo_elem(xyz, e2n[e], M)
        f = fib[e]
        D = np.outer(f, f) / vel[e, 0] ** 2 + (np.eye(3) - np.outer(f, f)) / vel[e, 1] ** 2
        x = xyz[e2n[e]]
        direct = [(x[i] - x[j]) @ D @ (x[i] - x[j]) for i, j in ((0, 1), (0, 2), (1, 2), (0, 3), (1, 3), (2, 3))]
        assert np.allclose(g, direct, rtol=1e-12)


def test_error_paths_return_codes():
    with pytest.raises(F.FimError, match="FIM_EIO"):
        F.Mesh.load_vtk("/nonexistent.vtk")
    with pytest.raises(F.FimError, match="FIM_EINVAL"):
        F.Mesh.from_arrays(np.zeros((4, 3)), np.array([[0, 1, 2, 9]]))
    with pytest.raises(F.FimError, match="repeated node"):
        F.Mesh.from_arrays(np.eye(4, 3), np.array([[0, 1, 2, 2]]))
    m = F.Mesh.load_vtk(UTV)
    with pytest.raises(F.FimError, match="no partition tags"):
        F.Pack(m, F.PART_TAGS)
    with pytest.raises(F.FimError, match="non-integer token"):
        m.read_tags(f"{cases.GOLDEN}/ut_mesh/ut_mesh.plan.json")
    import tempfile
    with tempfile.NamedTemporaryFile("w", suffix=".tags") as f:
        f.write("1\n2\n3\n")
        f.flush()
        with pytest.raises(F.FimError, match="entries"):
            m.read_tags(f.name)


def test_cli_usage_and_arguments():
    cli = F.CLI_PATH
    r = subprocess.run([cli, "--help"], capture_output=True, text=True)
    assert r.returncode == 0 and "--mesh" in r.stdout
    r = subprocess.run([cli, "--reps=0", "--mesh=a", "--part=b", "--plan=c", "--odir=d"],
                       capture_output=True, text=True)
    assert r.returncode == 0 and "invalid number of reps!" in r.stdout
    r = subprocess.run([cli, "--bogus"], capture_output=True, text=True)
    assert r.returncode != 0
    r = subprocess.run([cli, "--mesh=/nonexistent.vtk", "--part=x", "--plan=y", "--odir=z"],
                       capture_output=True, text=True, cwd=ROOT)
    assert r.returncode != 0 and "file not found" in r.stderr


def test_mesh_concat_is_disjoint_union():
    from fimpy import workloads as W
    a = F.Mesh.gen_box(seed=1, **dict(cases.CUBE, nx=3, ny=3, nz=2))
    b = F.Mesh.gen_box(seed=2, **dict(cases.SLAB, nx=4, ny=3, nz=2))
    a.set_regions(W.quadrant_regions(a))
    b.set_regions(W.quadrant_regions(b))
    u = F.Mesh.concat([a, b])
    assert u.nn == a.nn + b.nn and u.ne == a.ne + b.ne
    xa, ea, _ = a.arrays()
    xb, eb, _ = b.arrays()
    xu, eu, _ = u.arrays()
    assert np.array_equal(xu, np.vstack([xa, xb]))
    assert np.array_equal(eu, np.vstack([ea, eb + a.nn]))
    fa, va = a.material()
    fb, vb = b.material()
    fu, vu = u.material()
    assert np.array_equal(fu, np.vstack([fa, fb])) and np.array_equal(vu, np.vstack([va, vb]))
    # a partition of the union never mixes components
    p = F.Pack(u, F.PART_RCB, max_nn=40, seed=1)
    e2s, _ = p.partition()
    assert not set(e2s[:a.ne]) & set(e2s[a.ne:])
    with pytest.raises(F.FimError, match="FIM_EINVAL"):
        F.Mesh.concat([])


def test_multi_mesh_workload_is_seeded():
    from fimpy import workloads as W
    parts = [(n, W.scaled(c, 8)) for n, c in W.MULTI]
    m, offs, cfgs = W.multi_mesh(parts, 1)
    s1 = W.multi_solves(cfgs, offs, 12, 3)
    s2 = W.multi_solves(cfgs, offs, 12, 3)
    for x, y in zip(s1, s2):
        assert x["part"] == y["part"] and np.array_equal(x["sites"], y["sites"])
        assert np.array_equal(x["scale"], y["scale"])
    bounds = offs + [m.nn]
    for s in s1:
        k = s["part"]
        assert np.all((s["sites"] >= bounds[k]) & (s["sites"] < bounds[k + 1]))
        assert np.all((s["scale"] >= 0.5) & (s["scale"] <= 1.5))


def test_shell_generator_is_seeded_conforming_and_scarred():
    from fimpy import workloads as W
    a = W.shell(1, h=2500.0)
    b = W.shell(1, h=2500.0)
    c = W.shell(2, h=2500.0)
    xa, ea, _ = a.arrays()
    xb, eb, _ = b.arrays()
    assert np.array_equal(xa, xb) and np.array_equal(ea, eb)
    fa, va = a.material()
    fc, vc = c.material()
    assert not np.array_equal(va, vc)  # scars move with the seed
    assert np.allclose(np.linalg.norm(fa, axis=1), 1.0)
    scar = va[:, 0] < 0.1
    assert 0.01 < scar.mean() < 0.5 and np.all(va[scar] == 0.05)
    assert np.all(va[~scar, 0] == 0.6) and np.all(va[~scar, 1] == 0.24)
    # element centres lie in the shell
    cen = xa[ea].mean(axis=1)
    s = W.SHELL
    ro = np.sqrt((cen[:, 0] / s["ao"]) ** 2 + (cen[:, 1] ** 2 + cen[:, 2] ** 2) / s["bo"] ** 2)
    ri = np.sqrt((cen[:, 0] / s["ai"]) ** 2 + (cen[:, 1] ** 2 + cen[:, 2] ** 2) / s["bi"] ** 2)
    assert ro.max() <= 1.0 + 1e-9 + 2500.0 / s["bo"] and ri.min() >= 1.0 - 2500.0 / s["bi"]
    # conforming: no face shared by more than two elements
    faces = np.sort(np.concatenate([ea[:, [1, 2, 3]], ea[:, [0, 2, 3]], ea[:, [0, 1, 3]], ea[:, [0, 1, 2]]]), axis=1)
    _, cnt = np.unique(faces, axis=0, return_counts=True)
    assert cnt.max() == 2
    # fibres tangent to the ellipsoidal shells (orthogonal to the normal,
    # which the generator evaluates at the hex centre, not the tet centroid)
    n = np.stack([cen[:, 0] / s["ao"] ** 2, cen[:, 1] / s["bo"] ** 2, cen[:, 2] / s["bo"] ** 2], 1)
    n /= np.linalg.norm(n, axis=1)[:, None]
    assert np.abs(np.einsum("ij,ij->i", n, fa)).max() < 0.1  # normal taken at the hex centre
    sites = W.inner_surface_sites(a, 3, 7)
    assert len(set(sites.tolist())) == 3


def test_partition_tags_roundtrip(tmp_path):
    """an own partition written as a tags file and read back as a tag
    partition gives the same submeshes (the reference's --part input)"""
    m = F.Mesh.gen_box(seed=3, **cases.scaled(cases.SLAB, 8))
    a = F.Pack(m, F.PART_RCB, max_nn=200, seed=4)
    path = str(tmp_path / "rcb.tags")
    a.save_tags(path)
    ea, _ = a.partition()
    assert np.array_equal(np.loadtxt(path, dtype=np.int64), ea)
    m.read_tags(path)
    b = F.Pack(m, F.PART_TAGS)
    eb, _ = b.partition()
    # tag partitions number submeshes in first-seen tag order (reference
    # numbering); the grouping is the same
    assert len(np.unique(eb)) == len(np.unique(ea))
    pairs = set(zip(ea.tolist(), eb.tolist()))
    assert len(pairs) == len(np.unique(ea))


def test_plan_speeds_by_region(tmp_path):
    import json
    m = F.Mesh.load_vtk(UTV)
    assert m.apply_plan_speeds(f"{cases.UT}/ut_mesh.plan.json") == m.ne
    _, vel = m.material()
    assert np.all(vel == 0.6)
    plan = json.load(open(f"{cases.UT}/ut_mesh.plan.json"))
    f2 = json.loads(json.dumps(plan["functions"]["fnorm"]))
    f2["CV"]["CVref"].update(vf=0.9, vs=0.3)
    plan["functions"]["fslow"] = f2
    plan["config"]["norm"]["tags"] = [1, 2]
    plan["config"]["slow"] = {"func": "fslow", "tags": [3, 4]}
    p = tmp_path / "p.json"
    p.write_text(json.dumps(plan))
    m2 = F.Mesh.load_vtk(UTV)
    assert m2.apply_plan_speeds(str(p)) == m2.ne
    _, _, tags = m2.arrays()
    _, vel2 = m2.material()
    slow = np.isin(tags, [3, 4])
    assert np.all(vel2[slow] == [0.9, 0.3]) and np.all(vel2[~slow] == 0.6)


def test_submesh_size_limits_are_refused():
    """local ids are 16-bit: a partition leaf above 65535 elements is refused
    when the pack is built, before anything reaches a device"""
    m = F.Mesh.gen_box(seed=1, **cases.scaled(cases.SLAB, 2))
    with pytest.raises(F.FimError, match="FIM_ELIMIT"):
        F.Pack(m, F.PART_RCB, max_nn=60000, seed=1)


def test_autotune_search_logic_with_a_stand_in_timer(tmp_path, monkeypatch):
    """Coordinate descent over the space with the timing replaced by a known
    function of the configuration: the tuner must end at the known optimum,
    report it as the best measured trial, skip register budgets that select
    no kernel variant (workgroups other than 256), and serve the cache."""
    import json
    import fimpy.autotune as A
    calls = []

    def rate(cfg):  # optimum: 1536 copies, workgroups of 256 on 5 waves, 2 sweeps, 1 ms,
        # units of 3, two contexts
        return (10000.0 - abs(cfg["max_nn"] - 1536) - 100 * abs(cfg["local_iters"] - 2)
                - abs(cfg["window"] - 1000.0) + (50 if cfg["block"] == 256 else 0)
                + (20 if cfg["block"] == 256 and cfg["waves"] == 5 else 0)
                - 7 * abs(cfg["unit"] - 3) + (30 if cfg["contexts"] == 2 else 0))

    def fake_measure(mesh, packs, cfg, seeds, scales, seed):
        calls.append(dict(cfg))
        return rate(cfg)

    monkeypatch.setattr(A, "_measure", fake_measure)
    monkeypatch.setattr(A, "signature", lambda mesh, n: f"stand-in/{n}")
    cache = str(tmp_path / "tune.json")
    best, trials = A.tune(None, [[0]] * 8, cache=cache, rounds=3)
    assert best == {"max_nn": 1536, "block": 256, "local_iters": 2, "window": 1000.0, "waves": 5,
                    "unit": 3, "contexts": 2}
    rates = {json.dumps(c, sort_keys=True): r for c, r in trials}
    assert rates[json.dumps(best, sort_keys=True)] == max(rates.values())
    assert all(c["waves"] == 0 for c in calls if c["block"] != 256)
    assert len(rates) == len(calls)  # no configuration measured twice
    n = len(calls)
    again, _ = A.tune(None, [[0]] * 8, cache=cache)
    assert again == best and len(calls) == n


def test_stream_driver_bookkeeping_with_a_stand_in_solver():
    """fimpy.stream.stream_solve against a stand-in solver whose solves take
    a known number of iterations: every solve is reported exactly once, from
    the instance it was seeded into, only after its iterations have run, and
    a finished instance is recycled before it is seeded again."""
    from fimpy.stream import stream_solve

    class StandIn:
        n_instances = 3

        def __init__(self):
            self.left = [0] * self.n_instances   # iterations still to run
            self.seeded = [None] * self.n_instances
            self.log = []
            self.clock = 0

        def reset(self):
            self.log.append("reset")

        def set_scaling(self, sc, inst):
            self.log.append(("scale", inst))

        def seed_batch(self, sets, t0, insts):
            for s, i in zip(sets, insts):
                assert self.left[i] == 0 and self.seeded[i] is None
                self.seeded[i] = int(s[0])
                self.left[i] = 1 + int(s[0]) % 7   # uneven solve lengths

        def advance(self, iters, reset_stats=False):
            # like fim_advance: whole captured graphs of 6 iterations, counted
            # by the device from the last reset_stats call
            run = -(-iters // 6) * 6
            self.clock = (0 if reset_stats else self.clock) + run
            self.left = [max(0, v - run) for v in self.left]

        def stats(self):
            return {"global_iters": self.clock}

        def active(self):
            return [1 if v > 0 else 0 for v in self.left]

        def recycle(self, insts):
            for i in insts:
                assert self.left[i] == 0
                self.seeded[i] = None

    s = StandIn()
    sets = [np.array([k]) for k in range(11)]
    got = []

    def on_map(k, inst):
        assert s.seeded[inst] == k and s.left[inst] == 0
        got.append(k)

    iters = stream_solve(s, sets, scales=[[1.0] * 4] * len(sets), on_map=on_map, check_iters=2)
    assert sorted(got) == list(range(len(sets)))
    assert iters == s.clock > 0 and iters % 6 == 0  # the iterations actually run
    with pytest.raises(RuntimeError, match="iteration cap"):
        stream_solve(StandIn(), sets, on_map=None, check_iters=2, max_iters=12)
    assert s.log[0] == "reset" and sum(1 for e in s.log if e != "reset") == len(sets)


ADAPTER = os.path.join(ROOT, "oracle", "_ref", "adapter_run")


@pytest.mark.skipif(not (os.path.isdir("/root/reference/src") and os.path.exists(ADAPTER)),
                    reason="needs the reference sources and oracle/_ref/adapter_run")
def test_integration_adapter_runs_from_the_reference_side():
    """INTEGRATION.md's CufimSolver, compiled against the reference headers
    and host code (`make -C oracle adapter`), takes the reference's mesh,
    tags, options and plan through the C ABI: mesh and partition are built,
    and without a GPU it stops with the device error at fim_create"""
    U = cases.UT
    r = subprocess.run([ADAPTER, f"{U}/ut_mesh.vtk", f"{U}/ut_mesh.tags",
                        os.path.join(ROOT, "data", "default.cfg"), f"{U}/ut_mesh.plan.json",
                        os.devnull], capture_output=True, text=True, timeout=60)
    out = r.stdout + r.stderr
    assert "num_stims: 1" in out  # the reference's plan parser ran
    assert r.returncode == 0 or "hipSetDevice" in out, out



def test_scheduled_windows_and_vertex_sample():
    """per-instance windows: the batch runs quantile(tmax) / window
    iterations and solve j gets tmax_j / K (never below the floor); the
    vertex sample holds the bounding-box corners"""
    from fimpy import warmstart as WS
    tmax = np.array([100e3, 200e3, 300e3, 400e3, 500e3])
    w = WS.scheduled_windows(tmax, 2000.0, quantile=0.5)
    K = 300e3 / 2000.0
    assert np.allclose(w, tmax / K)
    assert np.isclose(w[2], 2000.0)  # the median solve keeps the base window
    w = WS.scheduled_windows(np.array([1e3, 400e3]), 2000.0, quantile=1.0, floor=250.0)
    assert w[0] == 250.0 and np.isclose(w[1], 2000.0)
    m = F.Mesh.gen_box(seed=1, **dict(cases.CUBE, nx=4, ny=4, nz=4))
    sample = WS.window_sample(m, k=10, seed=3)
    xyz, _, _ = m.arrays()
    lo, hi = xyz.min(axis=0), xyz.max(axis=0)
    for c in ((lo[0], lo[1], lo[2]), (hi[0], hi[1], hi[2]), (lo[0], hi[1], lo[2])):
        nearest = int(np.argmin(np.sum((xyz - np.array(c)) ** 2, axis=1)))
        assert nearest in sample
    assert len(np.unique(sample)) == len(sample) <= 18
    assert np.array_equal(sample, WS.window_sample(m, k=10, seed=3))  # seeded


def test_solver_group_maps_instances_to_contexts():
    """SolverGroup's instance bookkeeping (host side only): instance i of the
    group is instance i - off[k] of context k; parts() gives each context
    its share of a batch of n solves"""
    import fimpy as F

    class Ctx:
        def __init__(self, n):
            self.n_instances, self.window, self.got = n, 2000.0, []

        def download(self, inst=0):
            self.got.append(inst)
            return (inst, None)

    a, b, c = Ctx(5), Ctx(4), Ctx(3)
    g = F.SolverGroup([a, b, c])
    assert g.n_instances == 12 and g.window == 2000.0
    assert [(x is y, lo, hi) for (x, lo, hi), y in zip(g.parts(12), (a, b, c))] == \
        [(True, 0, 5), (True, 5, 9), (True, 9, 12)]
    assert [(lo, hi) for _, lo, hi in g.parts(7)] == [(0, 5), (5, 7)]
    assert [(lo, hi) for _, lo, hi in g.parts(3)] == [(0, 3)]
    for i in range(12):
        g.download(inst=i)
    assert a.got == [0, 1, 2, 3, 4] and b.got == [0, 1, 2, 3] and c.got == [0, 1, 2]
    for bad in (-1, 12):
        with pytest.raises(F.FimError):
            g.download(inst=bad)
