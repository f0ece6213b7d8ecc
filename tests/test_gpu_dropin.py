"""The drop-in path on the GPU: fim_run_reference (tw::run, src/main.cpp:84-104
-> CudaSolverFIM::solve, src/eikonal/cuda_fim.cu:2386-2503) and the `cufim`
CLI as a subprocess, on the reference's ut_mesh with its plan, tags and
data/default.cfg, against the reference's own output data/output/cuda_fim_res.vtk.

The golden was written after 3 windows of 150 ms (its state census fixes
that), while ut_mesh.plan.json asks for 4 ("timesteps"), so the plan is
copied with timesteps = 3.  Tolerances: everything before POINT_DATA byte for
byte; states exact; phi to the file's six printed decimals (5e-7 plus the
rounding of the last printed digit); the map of the library path bit for bit
equal to the oracle's restatement of the same solve."""
import json
import os
import subprocess

import numpy as np
import pytest

import cases
import fimpy as F
import pyoracle as O
import vtkio

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CFG = os.path.join(ROOT, "data", "default.cfg")
MESH = f"{cases.UT}/ut_mesh.vtk"
TAGS = f"{cases.UT}/ut_mesh.tags"
GOLD = f"{cases.UT}/cuda_fim_res.vtk"


def _plan(tmp_path, steps=3, edit=None):
    plan = json.load(open(f"{cases.UT}/ut_mesh.plan.json"))
    plan["simulation"]["timesteps"] = steps
    if edit:
        edit(plan)
    p = tmp_path / f"plan{steps}.json"
    p.write_text(json.dumps(plan, indent=2))
    return str(p)


def _split(path):
    text = open(path).read()
    k = text.index("POINT_DATA")
    return text[:k], vtkio.read_vtk(path)["point_data"]


def _check_against_golden(path):
    head, pd = _split(path)
    ghead, gpd = _split(GOLD)
    assert head == ghead  # points, cells, cell types, element tags
    assert np.array_equal(pd["states"], gpd["states"])
    assert np.max(np.abs(pd["phi"] - gpd["phi"])) <= 5.0001e-7


def _oracle_map(steps):
    g = vtkio.read_vtk(MESH)
    om = O.Mesh(g["xyz"], g["cells"], atags=np.loadtxt(TAGS, dtype=np.int32))
    return O.compat_solve(om, cases.ut_plan_stimuli(), 150e3, steps)


def test_run_reference_writes_the_golden(tmp_path):
    out = tmp_path / "out.vtk"
    avg = F.run_reference(MESH, TAGS, _plan(tmp_path), str(out), reps=2, cfg=CFG)
    assert avg > 0.0
    _check_against_golden(str(out))
    # the written values are the oracle's map, printed
    ophi, ost, _ = _oracle_map(3)
    ref = tmp_path / "oracle.vtk"
    m = F.Mesh.load_vtk(MESH)
    m.save_vtk(str(ref), ophi, ost)
    assert out.read_bytes() == ref.read_bytes()


def test_cli_writes_the_golden(tmp_path):
    out = tmp_path / "cli.vtk"
    r = subprocess.run([F.CLI_PATH, f"--mesh={MESH}", f"--part={TAGS}", f"--plan={_plan(tmp_path)}",
                        f"--odir={out}", "--reps=2"], capture_output=True, text=True, cwd=ROOT,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    assert "run  1 took" in r.stdout and "run  2 took" in r.stdout
    assert "Average execution time for 2 runs:" in r.stdout
    _check_against_golden(str(out))
    # same bytes as the library entry point
    lib_out = tmp_path / "lib.vtk"
    F.run_reference(MESH, TAGS, _plan(tmp_path), str(lib_out), reps=1, cfg=CFG)
    assert out.read_bytes() == lib_out.read_bytes()


def test_cli_unit_test_configuration(tmp_path):
    """`cufim --test` runs the reference's unit-test configuration: ut_mesh
    with the plan as shipped (4 windows); the map is the oracle's"""
    out = tmp_path / "ut.vtk"
    r = subprocess.run([F.CLI_PATH, "--test", f"--odir={out}"], capture_output=True, text=True,
                       cwd=ROOT, timeout=120)
    assert r.returncode == 0, r.stderr
    ophi, ost, _ = _oracle_map(4)
    ref = tmp_path / "oracle.vtk"
    F.Mesh.load_vtk(MESH).save_vtk(str(ref), ophi, ost)
    assert out.read_bytes() == ref.read_bytes()


def test_event_order_and_the_first_miss(tmp_path):
    """The plan's stimuli are read in key order and inactive ones are dropped
    while parsing (io_json.cpp:154-160), so an inactive stimulus ordered
    before S0 changes nothing.  update_events_gpu then seeds the events in
    order at cur_time 0 and stops at the first one that does not start at or
    after it (cuda_fim.cu:636-640): a stimulus with a negative start ordered
    before S0 leaves every vertex far and unreached."""
    def inactive_first(plan):
        st = plan["simulation"]["stimuli"]
        st["A_off"] = dict(st["S0"], active=0)

    out = tmp_path / "off.vtk"
    F.run_reference(MESH, TAGS, _plan(tmp_path, edit=inactive_first), str(out), reps=1, cfg=CFG)
    _check_against_golden(str(out))

    def negative_first(plan):
        st = plan["simulation"]["stimuli"]
        st["A_neg"] = json.loads(json.dumps(st["S0"]))
        st["A_neg"]["protocol"]["start"] = -1.0

    out2 = tmp_path / "neg.vtk"
    F.run_reference(MESH, TAGS, _plan(tmp_path, edit=negative_first), str(out2), reps=1, cfg=CFG)
    pd = vtkio.read_vtk(str(out2))["point_data"]
    assert np.all(pd["states"] == F.FAR) and np.all(pd["phi"] == -1.0)


ADAPTER = os.path.join(ROOT, "oracle", "_ref", "adapter_run")


@pytest.mark.skipif(not os.path.exists(ADAPTER), reason="oracle/_ref/adapter_run not built "
                    "(needs the reference's host sources at build time)")
def test_reference_side_adapter_writes_the_golden(tmp_path):
    """the reference's own tw::run, mesh loader, connectivity and plan
    parser (compiled from its sources) with the CufimSolver adapter of
    INTEGRATION.md in place of CudaSolverFIM: the drop-in seen from the
    reference side"""
    out = tmp_path / "adapter.vtk"
    r = subprocess.run([ADAPTER, MESH, TAGS, CFG, _plan(tmp_path), str(out), "2"], capture_output=True,
                       text=True, cwd=ROOT, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    _check_against_golden(str(out))
    lib_out = tmp_path / "lib.vtk"
    F.run_reference(MESH, TAGS, _plan(tmp_path), str(lib_out), reps=1, cfg=CFG)
    assert out.read_bytes() == lib_out.read_bytes()


def test_run_reference_errors_are_reported():
    with pytest.raises(F.FimError, match="FIM_EIO"):
        F.run_reference("/nonexistent.vtk", TAGS, f"{cases.UT}/ut_mesh.plan.json", "", cfg=CFG)
    with pytest.raises(F.FimError):
        F.run_reference(MESH, TAGS, f"{cases.UT}/ut_mesh.plan.json", "", reps=0, cfg=CFG)
