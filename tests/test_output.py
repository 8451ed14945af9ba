"""Output files in the reference's formats and the importer (host code of
libsimbuca_hip.so; no GPU).

Pinned by the reference's own IonCloud writer: tests/golden/output_n7/ holds
the files it wrote (PrintParticles, PrintCloud, Delete; bin/ioncloud.cpp) for
the state in output_n7.npz, in both layouts.  The writer must reproduce every
data line byte for byte, including the reference's energies (long-double
masses) and PrintCloud's quirks (per-species sums divided by the total ion
count, vz*vy in the energy).  Only lines that carry wall-clock times and the
banner text may differ.  Where the reference build is present, the same
comparison runs live on fresh inputs.

The importer reads the last print back: exactly the printed numbers from a
6-digit file, and the state to within one ulp from a 17-digit file.
"""
import os

import numpy as np
import pytest

import oracle as O
import simbuca_amd as S
from conftest import load_golden

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden", "output_n7")
TIME_LINES = ("#Simulation started", "#simulation ended", "#End time")


def _comparable(text):
    """Lines that must match: everything but the wall-clock lines and the
    four banner lines after the first (free text in both writers)."""
    lines = text.splitlines()
    keep = []
    for i, line in enumerate(lines):
        if line.startswith(TIME_LINES) or (1 <= i <= 4 and line.startswith("#")):
            continue
        keep.append(line)
    return keep


def _write_like_reference(prefix, amu, pos, vel, lifetime, prints, per_particle, digits=6):
    """The sequence of the reference harness: print, then `prints` times
    pos += 1e-7 vel, vel *= 0.75, lifetime += 1e-7, print; then close."""
    w = S.Writer(prefix, ["%gu" % m for m in amu], amu, per_particle_files=per_particle,
                 digits=digits)
    t = float(lifetime)
    pos = np.array(pos, dtype=np.float64); vel = np.array(vel, dtype=np.float64)
    w.particles(t, pos, vel); w.cloud(t, pos, vel)
    for _ in range(prints):
        pos = pos + vel * 1e-7
        vel = vel * 0.75
        t += 1e-7
        w.particles(t, pos, vel); w.cloud(t, pos, vel)
    w.close(t, pos, vel)
    return pos, vel, t


@pytest.mark.parametrize("layout", ["pAll", "per_particle"])
def test_writer_matches_reference_files(tmp_path, layout):
    g = load_golden("output_n7.npz")
    _write_like_reference(str(tmp_path / "ref"), g["mass_amu"], g["pos"], g["vel"],
                          float(g["lifetime"]), int(g["prints"]), layout == "per_particle")
    want = sorted(os.listdir(os.path.join(GOLDEN, layout)))
    got = sorted(f for f in os.listdir(tmp_path) if f.startswith("ref_"))
    assert got == want
    for name in want:
        with open(os.path.join(GOLDEN, layout, name)) as fh:
            ref = _comparable(fh.read())
        with open(tmp_path / name) as fh:
            mine = _comparable(fh.read())
        assert mine == ref, name
        assert any(not line.startswith("#") for line in ref)      # data lines were compared


@pytest.mark.skipif(not O.ref_available(), reason="reference build (oracle/_ref) not present")
def test_writer_matches_live_reference(tmp_path, rng):
    n = 40
    amu = rng.choice(np.array([133.0, 85.0, 39.0, 7.0]), n)
    ch = np.ones(n, dtype=np.int32)
    pos = rng.normal(0, 2e-3, (n, 3)); vel = rng.normal(0, 300.0, (n, 3))
    ref_dir = tmp_path / "ref"; ref_dir.mkdir()
    files = O.ref_output(amu, ch, pos, vel, 1.25e-5, 3, str(ref_dir))
    _write_like_reference(str(tmp_path / "ref"), amu, pos, vel, 1.25e-5, 3, False)
    for name, text in files.items():
        with open(tmp_path / ("ref" + name)) as fh:
            assert _comparable(fh.read()) == _comparable(text), name


def test_import_reads_the_last_print(tmp_path):
    g = load_golden("output_n7.npz")
    # the reference's own file: 6 significant digits
    names, pos, vel, t = S.import_snapshot(os.path.join(GOLDEN, "pAll", "ref"))
    fpos, fvel, ft = _write_like_reference(str(tmp_path / "x"), g["mass_amu"], g["pos"], g["vel"],
                                           float(g["lifetime"]), int(g["prints"]), False)
    assert names == ["%gu" % m for m in g["mass_amu"]]
    assert t == pytest.approx(ft, rel=1e-6)
    assert np.allclose(pos, fpos, rtol=1e-6, atol=0)
    assert np.allclose(vel, fvel, rtol=1e-6, atol=0)
    # the printed numbers themselves come back exactly
    with open(os.path.join(GOLDEN, "pAll", "ref_pAll.txt")) as fh:
        last = [ln.split() for ln in fh if ln[0].isdigit()][-len(names):]
    assert np.array_equal(pos, np.array([[float(c) / 1000. for c in r[3:6]] for r in last]))
    assert np.array_equal(vel, np.array([[float(c) for c in r[6:9]] for r in last]))


def test_import_round_trip_with_17_digits(tmp_path, rng):
    n = 33
    amu = rng.choice(np.array([133.0, 85.0]), n)
    pos = rng.normal(0, 1e-3, (n, 3)); vel = rng.normal(0, 137.0, (n, 3))
    fpos, fvel, ft = _write_like_reference(str(tmp_path / "c"), amu, pos, vel, 3.3e-6, 1, False,
                                           digits=17)
    names, ipos, ivel, t = S.import_snapshot(str(tmp_path / "c"))
    assert len(names) == n
    assert np.array_equal(ivel, fvel)                                # m/s: printed exactly
    ulp = np.spacing(np.abs(fpos))
    assert np.all(np.abs(ipos - fpos) <= ulp)                        # through mm: one ulp
    assert abs(t - ft) <= np.spacing(ft)


def test_import_errors(tmp_path):
    with pytest.raises(S.SimbucaError, match="cannot open"):
        S.import_snapshot(str(tmp_path / "missing"))
    (tmp_path / "bad_pAll.txt").write_text("# header\n0 133u 1.0 2.0 3.0\n")
    with pytest.raises(S.SimbucaError, match="11 columns"):
        S.import_snapshot(str(tmp_path / "bad"))
    (tmp_path / "big_pAll.txt").write_text(
        "".join(f"{i} 133u 0 0 0 0 0 0 0 0\n" for i in range(5)))
    with pytest.raises(S.SimbucaError, match="capacity"):
        S.import_snapshot(str(tmp_path / "big"), capacity=4)
    with pytest.raises(S.SimbucaError, match="one non-empty word"):
        S.Writer(str(tmp_path / "w"), ["two words"], [133.0])
