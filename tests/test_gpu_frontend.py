"""Simulation files run on the GPU (sbk_run_ini) against the files the
reference program itself wrote for them (tests/golden/ini/<case>/, made by
tests/golden/make_golden.py from oracle/_ref/simbuca), and the engine calls
the runner relies on: sbk_advance_until and the t_prev status field.  The
cases cover every operation type (normal, trap_ramp, tof, nop, dissociation
before and after the first move), fixed and adaptive steps, Coulomb on and
off, both file layouts, print_after_operation and the print_at_x detector
plane.

Tolerance of the file comparison: the reference prints "%.6e"; FP64-exact
runs agree with it far below that, so two prints of the same number differ
by at most one unit in the 7th significant digit.  Each number must agree to
3e-6 of itself plus 3e-6 of its column's largest magnitude in that file (for
entries near zero); indices, names, ion counts and line counts must be
identical, so the print schedule is the reference's.
"""
import os
import shutil
import subprocess

import numpy as np
import pytest

import simbuca_amd as S
from test_gpu_engine import _rand_cloud

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
INI = os.path.join(HERE, "golden", "ini")
CASES = ["paul_cycle", "adaptive_separate", "two_ramps", "tof_detector", "adaptive_ramp"]


def _tables(root):
    """{relative path: [data rows as word lists]} of the .txt output files
    under root (the log file is compared by _logs)."""
    out = {}
    for d, _, files in os.walk(root):
        for f in files:
            if f.endswith(".txt") and not f.endswith("_logfile.txt"):
                path = os.path.join(d, f)
                with open(path) as fh:
                    out[os.path.relpath(path, root)] = [ln.split() for ln in fh if ln[:1].isdigit()]
    return out


def _logs(root):
    """{relative path: lines} of the log files under root, without the lines
    that carry wall-clock time."""
    out = {}
    for d, _, files in os.walk(root):
        for f in files:
            if f.endswith("_logfile.txt"):
                path = os.path.join(d, f)
                with open(path) as fh:
                    out[os.path.relpath(path, root)] = [
                        ln for ln in fh.read().splitlines()
                        if not ln.startswith(("logfile created on", "#simulation ended after"))]
    return out


def _compare(mine, ref, rtol=3e-6, what=""):
    assert len(mine) == len(ref), what
    # integer and name columns: the first 2 (particle and X-plane files) or 3 (cloud files)
    n_words = 3 if "_cloud" in what else 2
    for a, b in zip(mine, ref):
        assert len(a) == len(b) and a[:n_words] == b[:n_words], what
    A = np.array([[float(x) for x in r[n_words:]] for r in mine])
    B = np.array([[float(x) for x in r[n_words:]] for r in ref])
    scale = np.abs(B).max(axis=0, keepdims=True)
    bad = np.abs(A - B) > rtol * np.abs(B) + rtol * scale
    assert not bad.any(), (what, np.argwhere(bad)[:5], A[bad][:5], B[bad][:5])


def _run(case, tmp_path, monkeypatch, precision):
    print(f"run_ini {case} precision {precision}", flush=True)
    run = tmp_path / "run"
    run.mkdir()
    shutil.copy(os.path.join(INI, case + ".ini"), run / "sim.ini")
    monkeypatch.chdir(run)
    S.run_ini("sim.ini", precision=precision, time_mode=S.TIME_FROZEN)
    return _tables(str(run)), _tables(os.path.join(INI, case))


@pytest.mark.parametrize("case", CASES)
def test_run_ini_matches_reference_program(case, tmp_path, monkeypatch):
    mine, ref = _run(case, tmp_path, monkeypatch, S.FP64_EXACT)
    assert sorted(mine) == sorted(ref)
    for name in ref:
        _compare(mine[name], ref[name], what=name)
    # the log file, line for line (wall-clock lines aside)
    ref_logs = _logs(os.path.join(INI, case))
    assert ref_logs and _logs(str(tmp_path / "run")) == ref_logs


@pytest.mark.parametrize("precision,rtol", [(S.FP64, 1e-5), (S.MIXED, 1e-4), (S.MIXED_KAHAN, 1e-4)])
def test_run_ini_other_precisions(tmp_path, monkeypatch, precision, rtol):
    # same schedule and names; the other Coulomb sums move the numbers by far
    # less than the cloud's size over this short run (tiled FP64 differs from
    # the reference's order in the last bits, mixed by about 1e-7)
    mine, ref = _run("paul_cycle", tmp_path, monkeypatch, precision)
    assert sorted(mine) == sorted(ref)
    for name in ref:
        _compare(mine[name], ref[name], rtol=rtol, what=name)


SPLIT = """[cloud]
NParticles = 10
temp = 0.01
sx = 0.0005
sy = 0.0005
sz = 0.001
[composition]
133Xx+ = 0.5
40Zz2+ = 0.5
[ode]
method = DP
timestep = 1e-8
[coulomb]
enable = true
[trap]
Vrf = 50
Vdc = 5
r0 = 0.004
z0 = 0.004
freq_rf = 1e6
[output]
prefix = {prefix}
timestep = 1e-7
digits = 17
"""
OPS = ["[operation0]\ntype = normal\nduration = 4.55e-7\n",
       "[operation1]\ntype = tof\nduration = 3.35e-7\nEkick = 100\n",
       "[operation2]\ntype = nop\nduration = 2.05e-7\n"]


def _last_print(prefix, n):
    names, pos, vel, t = S.import_snapshot(prefix)
    assert len(names) == n
    return names, pos, vel, t


def test_run_ini_import_continues_a_run(tmp_path, monkeypatch):
    """A run split in two by [import] of a 17-digit print ends where the whole
    run ends (the print carries positions to within an ulp)."""
    print("run_ini import split", flush=True)
    monkeypatch.chdir(tmp_path)
    (tmp_path / "whole.ini").write_text(SPLIT.format(prefix="whole") + "".join(OPS))
    (tmp_path / "a.ini").write_text(SPLIT.format(prefix="a") + OPS[0])
    b = SPLIT.format(prefix="b").replace(SPLIT[:SPLIT.index("[ode]")], "[import]\nprefix = a\n")
    (tmp_path / "b.ini").write_text(b + "".join(OPS[1:]))
    for name in ("whole", "a", "b"):
        S.run_ini(name + ".ini", precision=S.FP64_EXACT)
    nw, pw, vw, tw = _last_print("whole", 10)
    na, pa, va, ta = _last_print("a", 10)
    nb, pb, vb, tb = _last_print("b", 10)
    assert nw == na == nb and nw[0] == "133Xx+" and nw[-1] == "40Zz2+"
    assert ta == pytest.approx(4.6e-7, rel=1e-12)
    assert tb == pytest.approx(tw, rel=1e-14)
    assert np.allclose(pb, pw, rtol=1e-9, atol=1e-9 * np.abs(pw).max())
    assert np.allclose(vb, vw, rtol=1e-9, atol=1e-9 * np.abs(vw).max())
    d = S.describe_ini("b.ini")
    assert d["import"] == "a" and d["lifetime"] == ta and d["N"] == 10


def test_program_runs_a_simulation_file(tmp_path):
    """simbuca_hip, the command-line program, as `simbuca sim.ini` is run."""
    exe = os.path.join(os.path.dirname(HERE), "simbuca_amd", "simbuca_hip")
    assert os.path.exists(exe), "build it with make -C simbuca_amd/csrc"
    run = tmp_path / "run"
    run.mkdir()
    shutil.copy(os.path.join(INI, "two_ramps.ini"), run / "sim.ini")
    print("simbuca_hip two_ramps", flush=True)
    p = subprocess.run([exe, "--precision", "fp64_exact", "sim.ini"], cwd=str(run),
                       capture_output=True, text=True, timeout=300)
    assert p.returncode == 0, p.stderr
    assert "operation trap_ramp" in p.stderr and "Program ended" in p.stderr
    mine, ref = _tables(str(run)), _tables(os.path.join(INI, "two_ramps"))
    assert sorted(mine) == sorted(ref)
    for name in ref:
        _compare(mine[name], ref[name], what=name)
    assert _logs(str(run)) == _logs(os.path.join(INI, "two_ramps"))
    p = subprocess.run([exe, "--precision", "fp16", "sim.ini"], cwd=str(run),
                       capture_output=True, text=True, timeout=60)
    assert p.returncode == 2 and "usage" in p.stderr


def test_run_ini_errors(tmp_path, monkeypatch):
    monkeypatch.chdir(tmp_path)
    with pytest.raises(S.SimbucaError, match="cannot read"):
        S.run_ini("missing.ini")
    (tmp_path / "x.ini").write_text("[cloud]\nNParticles = 4\n[composition]\n9Q+ = 1\n"
                                    "[ode]\nmethod = RK\n")
    with pytest.raises(S.SimbucaError, match="Dormand-Prince"):
        S.run_ini("x.ini")


def test_advance_until_and_t_prev(rng):
    amu, m, ch, pos, vel = _rand_cloud(rng, 64)
    f = S.default_fields(trap=1, Vrf=0.0, Vdc=5.0, kappa=0.001, coulomb=1)
    with S.Engine(precision="fp64_exact", adaptive=False, use_graph=False, h0=1e-9) as e:
        e.add_cloud(pos, vel, m, ch, f)
        e.advance_until(5.5e-9)                         # stops at the first step past it
        st = e.status(0)
        t = 0.0
        for _ in range(6):
            t_prev, t = t, t + 1e-9
        assert st["lifetime"] == t and st["t_prev"] == t_prev and st["steps"] == 6
        e.advance_until(5.5e-9)                         # already there: no step
        assert e.status(0)["steps"] == 6
        e.advance_until(np.nextafter(t, np.inf))        # exactly one step
        st = e.status(0)
        assert st["steps"] == 7 and st["t_prev"] == t and st["lifetime"] == t + 1e-9
        p1, v1 = e.state(0)
    # the same 7 steps through sbk_step
    with S.Engine(precision="fp64_exact", adaptive=False, use_graph=False, h0=1e-9) as e:
        e.add_cloud(pos, vel, m, ch, f)
        for _ in range(7):
            e.step()
        p2, v2 = e.state(0)
    assert np.array_equal(p1, p2) and np.array_equal(v1, v2)
    # adaptive: the end lies inside the last step
    with S.Engine(precision="fp64_exact", adaptive=True, use_graph=True, h0=1e-9) as e:
        e.add_cloud(pos, vel, m, ch, f)
        e.advance_until(3.3e-7)
        st = e.status(0)
        assert st["t_prev"] < 3.3e-7 <= st["lifetime"]
        assert st["steps"] > 1


def _host_crossings(p0, p1, v1, xp, steps, t):
    """The reference's per-step test (bin/ion_fly.cpp:110-114) on the host."""
    out = []
    for i in range(len(p0)):
        if (p0[i, 0] < xp and p1[i, 0] > xp) or (p0[i, 0] > xp and p1[i, 0] < xp):
            out.append((steps, i, t, p1[i].copy(), v1[i].copy()))
    return out


def _check_crossings(got, ref):
    assert len(ref) > 20, len(ref)
    assert list(got["step"]) == [r[0] for r in ref]
    assert list(got["ion"]) == [r[1] for r in ref]
    assert np.array_equal(got["t"], [r[2] for r in ref])
    assert np.array_equal(got["pos"], np.array([r[3] for r in ref]))
    assert np.array_equal(got["vel"], np.array([r[4] for r in ref]))


def test_crossings_match_a_step_by_step_check(rng):
    # FP64-exact Penning cloud: the device list after one advance over many
    # steps against the reference's test applied after each of the same steps
    amu, m, ch, pos, vel = _rand_cloud(rng, 400)
    f = S.default_fields(trap=0, coulomb=1, V0=10.0, d0=11e-3, Bz=7.0)
    xp = 2e-5
    with S.Engine(precision="fp64_exact", h0=2e-9) as e:
        e.add_cloud(pos, vel, m, ch, f)
        assert e.set_crossing_plane(0, xp) == 65536
        e.advance(1e-6)
        got = e.take_crossings()
        steps = e.status(0)["steps"]
        assert e.take_crossings()["ion"].size == 0          # taking empties the list
    ref = []
    with S.Engine(precision="fp64_exact", h0=2e-9) as e:
        e.add_cloud(pos, vel, m, ch, f)
        p0, _ = e.state(0)
        for _ in range(steps):
            e.step()
            p1, v1 = e.state(0)
            st = e.status(0)
            ref += _host_crossings(p0, p1, v1, xp, st["steps"], st["lifetime"])
            p0 = p1
    _check_crossings(got, ref)
    assert np.all(got["cloud"] == 0)


def test_crossings_survive_reordering_and_can_stop(rng):
    # mixed precision in spatial order: re-sorting between steps changes what
    # the device slots mean, so entries still on the list must be mapped to
    # the caller's ions first
    amu, m, ch, pos, vel = _rand_cloud(rng, 600)
    f = S.default_fields(trap=0, coulomb=1, V0=10.0, d0=11e-3, Bz=7.0)
    xp = -1e-5
    with S.Engine(precision="mixed", h0=2e-9, reorder_interval=0) as e:
        e.add_cloud(pos, vel, m, ch, f)
        e.set_crossing_plane(0, xp)
        ref = []
        p0, _ = e.state(0)
        for k in range(150):
            e.step()
            p1, v1 = e.state(0)
            st = e.status(0)
            ref += _host_crossings(p0, p1, v1, xp, st["steps"], st["lifetime"])
            p0 = p1
            if k % 7 == 3:
                e.reorder()
        _check_crossings(e.take_crossings(), ref)
        # enable = 0 stops recording
        e.set_crossing_plane(0, xp, enable=False)
        e.advance(5e-8)
        assert e.take_crossings()["ion"].size == 0


def test_crossings_of_two_clouds_and_a_short_buffer(rng):
    # two clouds with their own planes; entries come ordered by cloud, then
    # step, then ion; moving a plane takes effect from the next step; a
    # buffer smaller than the list is refused and the entries stay
    import ctypes
    clouds = [_rand_cloud(rng, n) for n in (300, 200)]
    f = S.default_fields(trap=0, coulomb=1, V0=10.0, d0=11e-3, Bz=7.0)
    planes = (1e-5, -2e-5)
    with S.Engine(precision="fp64_exact", h0=2e-9) as e:
        for amu, m, ch, pos, vel in clouds:
            e.add_cloud(pos, vel, m, ch, f)
        for c, xp in enumerate(planes):
            e.set_crossing_plane(c, xp)
        ref = {0: [], 1: []}
        p0 = [e.state(c)[0] for c in (0, 1)]
        for k in range(120):
            if k == 60:
                planes = (planes[0], 3e-5)
                e.set_crossing_plane(1, planes[1])
            e.step()
            for c in (0, 1):
                p1, v1 = e.state(c)
                st = e.status(c)
                ref[c] += _host_crossings(p0[c], p1, v1, planes[c], st["steps"], st["lifetime"])
                p0[c] = p1
        n = ctypes.c_int()
        L = S.lib()
        assert L.sbk_take_crossings(e._h, 0, ctypes.byref(n), None, None, None, None, None, None) == 0
        assert n.value == len(ref[0]) + len(ref[1]) > 20
        ion = np.zeros(n.value, np.int32)
        rc = L.sbk_take_crossings(e._h, n.value - 1, ctypes.byref(n), None, S._ip(ion), None, None,
                                  None, None)
        assert rc == -1                                 # SBK_EINVAL
        got = e.take_crossings()                        # still all there
    k0 = len(ref[0])
    assert list(got["cloud"]) == [0] * k0 + [1] * len(ref[1])
    sel0 = {key: v[:k0] for key, v in got.items()}
    sel1 = {key: v[k0:] for key, v in got.items()}
    for sel, r in ((sel0, ref[0]), (sel1, ref[1])):
        assert list(sel["step"]) == [x[0] for x in r] and list(sel["ion"]) == [x[1] for x in r]
        assert np.array_equal(sel["t"], [x[2] for x in r])
        assert np.array_equal(sel["pos"], np.array([x[3] for x in r]).reshape(-1, 3))


def test_empty_cloud_file_ends_and_writes_its_files(tmp_path, monkeypatch):
    """NParticles = 0: the move loop is skipped, as move_particles leaves at
    once for an empty cloud (bin/ion_fly.cpp:87-88); every operation, the
    prints after each and the closing block still run.  (The reference
    program itself ends with a segmentation fault on such a file, so there is
    no reference output to compare with.)"""
    print("run_ini empty cloud", flush=True)
    run = tmp_path / "run"
    run.mkdir()
    with open(os.path.join(INI, "paul_cycle.ini")) as fh:
        text = fh.read().replace("NParticles = 12", "NParticles = 0")
    assert "NParticles = 0" in text
    (run / "sim.ini").write_text(text)
    monkeypatch.chdir(run)
    S.run_ini("sim.ini", precision=S.FP64_EXACT, time_mode=S.TIME_FROZEN)
    out = run / "out"
    # no ion, so no particle line: the log is the only file, and it is closed
    assert sorted(os.listdir(out)) == ["paul_cycle_logfile.txt"]
    log = (out / "paul_cycle_logfile.txt").read_text().splitlines()
    # the lines the reference program logs for this file before its crash
    want = ["0 particles created. Max-Boltz distribution around 0.01 eV. ",
            "Normal trap operation for 1e-06 sec,", "Normal trap operation for 5e-07 sec,",
            "Normal trap operation for 4e-07 sec,", "Normal trap operation for 3e-07 sec,",
            "Close everything..."]
    got = [ln for ln in log if ln in want]
    assert got == want, log
    assert log[-1].startswith("#simulation ended after")
