"""The reference's simulation set-up in libsimbuca_hip.so (host code, no GPU):
InitCloud, the mass tables, the [composition] rules and the simulation-file
reader behind sbk_run_ini.

Pinned by the reference's own code:
  * tests/golden/init_cloud.npz holds clouds drawn by the reference's
    InitCloud (bin/initialization.cpp:117-206) for each of its energy
    branches; sbk_init_cloud must reproduce them bit for bit;
  * tests/golden/masses.json holds masses its IonTable/PDGTable give;
  * tests/golden/ini/<case>/ holds the files the reference program wrote for
    tests/golden/ini/<case>.ini; the ion names and species counts in them
    pin the composition rules;
  * where the reference build is present (oracle/_ref/ref_init, simbuca), the
    same comparisons run live on fresh inputs.
Closed forms fix what those leave open: the mean kinetic energy of the
reference's speed law is 3/2 E, directions are isotropic, positions fill the
ellipsoid.  The simulation-file rules (inih, bin/ini.c; INIReader,
bin/inireader.cpp) are checked case by case.
"""
import json
import os
import subprocess

import numpy as np
import pytest

import oracle as O
import simbuca_amd as S

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden")
INI = os.path.join(GOLDEN, "ini")
AME = os.path.join(O.REF_LIBRARIES, "ame", "mass.mas12")
PDG = os.path.join(O.REF_LIBRARIES, "pdg", "pdg_table.txt")
have_tables = pytest.mark.skipif(not (os.path.exists(AME) and os.path.exists(PDG)),
                                 reason="the reference's mass tables are not present")
have_refsim = pytest.mark.skipif(not O.refsim_available(),
                                 reason="reference program (oracle/_ref/simbuca, ref_init) not built")


# ---------------------------------------------------------------------------
# InitCloud
# ---------------------------------------------------------------------------

def test_init_cloud_matches_reference_golden():
    g = np.load(os.path.join(GOLDEN, "init_cloud.npz"))
    amu = g["mass_amu"]
    k = 0
    while f"seed{k}" in g:
        pos, vel = S.init_cloud(amu.size, amu, g[f"energy{k}"], g[f"semiaxis{k}"], g[f"offset{k}"],
                                int(g[f"seed{k}"]))
        assert np.array_equal(pos, g[f"pos{k}"]), k
        assert np.array_equal(vel, g[f"vel{k}"]), k
        k += 1
    assert k == 3      # the one-energy, two-energy and three-energy branches


@have_refsim
@pytest.mark.parametrize("seed,energy", [(1, (0.3,)), (2024, (0.02, 0.5)), (99, (1e-3,) * 3)])
def test_init_cloud_matches_live_reference(seed, energy, rng):
    n = 700
    amu = rng.choice(np.array([133.0, 40.0, 4.0026, 0.000548579909]), n)
    semi = rng.uniform(1e-4, 5e-3, 3)
    off = rng.normal(0, 1e-3, 3)
    rp, rv = O.ref_init_cloud(n, seed, energy, semi, off, amu)
    pos, vel = S.init_cloud(n, amu, energy, semi, off, seed)
    assert np.array_equal(pos, rp) and np.array_equal(vel, rv)


def test_init_cloud_closed_forms():
    n = 40000
    E = 0.05
    amu = np.full(n, 85.0)
    semi = np.array([1e-3, 2e-3, 5e-4]); off = np.array([1e-4, 0.0, -3e-4])
    pos, vel = S.init_cloud(n, amu, E, semi, off, seed=3)
    # positions fill the ellipsoid: inside it, with the uniform-ball moments
    r = (pos - off) / semi
    rho2 = (r * r).sum(1)
    assert rho2.max() <= 1.0 + 1e-12
    assert np.mean(rho2) == pytest.approx(3.0 / 5.0, rel=0.01)       # <r^2> of the unit ball
    assert np.abs(r.mean(0)).max() < 0.01
    # KE = -E log(1 - sqrt(u)), u uniform: mean 3/2 E
    ke = 0.5 * 85.0 * S.AMU * (vel * vel).sum(1) / S.EL_CHARGE
    assert ke.mean() == pytest.approx(1.5 * E, rel=0.02)
    # isotropic directions
    d = vel / np.linalg.norm(vel, axis=1)[:, None]
    assert np.abs(d.mean(0)).max() < 0.015
    assert np.allclose((d * d).mean(0), 1.0 / 3.0, atol=0.01)
    # the separate axial energy only changes |vz|
    _, v2 = S.init_cloud(n, amu, (E, 4 * E), semi, off, seed=3)
    ke_z = 0.5 * 85.0 * S.AMU * v2[:, 2] ** 2 / S.EL_CHARGE
    ke_z1 = 0.5 * 85.0 * S.AMU * vel[:, 2] ** 2 / S.EL_CHARGE
    assert ke_z.mean() == pytest.approx(4 * ke_z1.mean(), rel=0.1)


def test_init_cloud_first_draws_by_hand():
    """Seed 0, unit ball: the first point is (2u1-1, 2u2-1, 2u3-1) of the
    MT19937 outputs 2357136044, 2546248239, 3071714933 (the standard's seed-0
    sequence) when it falls inside the ball, which it does."""
    u = np.array([2357136044, 2546248239, 3071714933]) / 2.0 ** 32
    p = 2 * u - 1
    assert (p * p).sum() <= 1
    pos, _ = S.init_cloud(1, [133.0], 0.1, (1.0, 1.0, 1.0), (0.0, 0.0, 0.0), seed=0)
    assert np.array_equal(pos[0], p)


def test_init_cloud_errors():
    with pytest.raises(S.SimbucaError, match="bad arguments"):
        S.init_cloud(3, [1.0] * 3, (1.0, 2.0, 3.0, 4.0))


# ---------------------------------------------------------------------------
# mass tables and species names
# ---------------------------------------------------------------------------

@have_tables
def test_masses_match_reference_golden():
    t = S.MassTable(AME, PDG)
    rows = json.load(open(os.path.join(GOLDEN, "masses.json")))
    assert {r["table"] for r in rows} == {"ame", "pdg", "none"}
    for r in rows:
        if r["table"] == "none":
            with pytest.raises(S.SimbucaError, match="neither"):
                t.mass(r["name"])
        else:
            assert t.mass(r["name"]) == float(r["mass_amu"]), r["name"]


@have_tables
def test_masses_published_values():
    t = S.MassTable(AME, PDG)
    # AME2012 (mass.mas12) values in micro-u, and 12C by definition
    assert t.mass("12C") == 12.0
    assert t.mass("133Cs") == pytest.approx(132.905451961, abs=1e-12)
    assert t.mass("1H") == pytest.approx(1.00782503223, abs=1e-12)
    # PDG: m_p = 0.938272 GeV over 931.494 MeV/u; the antiparticle takes its particle's mass
    assert t.mass("proton") == pytest.approx(938.272 / 931.494, rel=1e-12)
    assert t.mass("e+") == t.mass("e-") == pytest.approx(0.51099891 / 931.494, rel=1e-12)
    # AME before PDG: "1n" is the AME neutron, "neutron" the PDG one
    assert t.mass("1n") != t.mass("neutron")
    assert t.mass("1n") == pytest.approx(t.mass("neutron"), rel=1e-6)


@have_refsim
def test_masses_match_live_reference():
    names = ["56Fe", "197Au", "9Be", "19F", "anti-proton", "proton", "K+", "Zz", "100Sn"]
    ref = O.ref_masses(names)
    t = S.MassTable(AME, PDG)
    for n in names:
        kind, m = ref[n]
        if kind == "none":
            with pytest.raises(S.SimbucaError):
                t.mass(n)
        else:
            assert t.mass(n) == m, n


@have_refsim
@have_tables
def test_masses_every_table_row_matches_reference():
    """Every name in both tables (AME names from the fixed columns, PDG names
    from the particle lines) looked up here and by the reference's own
    IonTable/PDGTable: the same table, the same mass, or absent from both."""
    names = set()
    for line in open(AME, encoding="latin-1"):
        line = line.rstrip("\n")
        if len(line) > 115 and len(line) != 124:
            line = line.replace("#", " ")
            el = line[20:22]
            el = el.replace(" ", "", 1)
            names.add("%d%s" % (int(line[16:19]), el))
    for line in open(PDG, encoding="latin-1"):
        w = line.split()
        if len(w) >= 3 and not line.startswith("#") and w[0].isdigit() and w[2].lstrip("-").isdigit():
            names.add(w[1])
    names = sorted(n for n in names if n and " " not in n)
    assert len(names) > 3000
    ref = {}
    for k in range(0, len(names), 500):
        ref.update(O.ref_masses(names[k:k + 500]))
    t = S.MassTable(AME, PDG)
    for n in names:
        kind, m = ref[n]
        if kind == "none":
            with pytest.raises(S.SimbucaError):
                t.mass(n)
        else:
            assert t.mass(n) == m, n


def test_species_names():
    # charge suffixes (bin/initialization.cpp:339-355) and the leading-digit mass
    assert S.parse_species("133Xx+") == (133.0, 1)
    assert S.parse_species("85Yy2+") == (85.0, 2)
    assert S.parse_species("40Zz-") == (40.0, -1)
    assert S.parse_species("7Ww3-") == (7.0, -3)
    assert S.parse_species("133Xx") == (133.0, 1)
    assert S.parse_species("12") == (12.0, 1)
    for bad in ("Xx+", "+", "-"):
        with pytest.raises(S.SimbucaError, match="no mass"):
            S.parse_species(bad)


@have_tables
def test_species_names_with_tables():
    t = S.MassTable(AME, PDG)
    assert S.parse_species("133Cs+", t) == (t.mass("133Cs"), 1)
    assert S.parse_species("40Ar2+", t) == (t.mass("40Ar"), 2)
    assert S.parse_species("e--", t) == (t.mass("e-"), -1)        # "e-" with charge -1
    assert S.parse_species("proton+", t) == (t.mass("proton"), 1)


def test_mass_table_errors(tmp_path):
    with pytest.raises(S.SimbucaError, match="cannot open"):
        S.MassTable(str(tmp_path / "missing.mas"))
    (tmp_path / "short.mas").write_text("header\n" * 5)
    with pytest.raises(S.SimbucaError, match="no nuclide rows"):
        S.MassTable(str(tmp_path / "short.mas"))


def test_mass_table_columns(tmp_path):
    """The fixed columns of IonTable (bin/IonTable.cpp:35-79) on rows written
    here: A at [16,19), El at [20,22) less its first blank, integer mass at
    [96,99), micro-u at [100,113); '#' reads as a blank; rows are the lines
    longer than 115 characters other than 124."""
    def row(A, el, whole, micro, pad=120):
        line = [" "] * pad
        line[16:19] = list("%3d" % A)
        line[20:22] = list("%-2s" % el)
        line[96:99] = list("%3d" % whole)
        line[100:113] = list("%13s" % micro)
        return "".join(line)
    rows = [row(133, "Xq", 132, "905451.961"), row(7, "Q", 7, "016003.43#"),
            row(9, "Zy", 9, "999999.000", pad=124)]       # 124 characters: skipped
    (tmp_path / "t.mas").write_text("\n".join(["x" * 200] + rows) + "\n")
    t = S.MassTable(str(tmp_path / "t.mas"))
    assert t.mass("133Xq") == 132 + 905451.961e-6
    assert t.mass("7Q") == 7 + 16003.43e-6
    with pytest.raises(S.SimbucaError):
        t.mass("9Zy")


# ---------------------------------------------------------------------------
# simulation files
# ---------------------------------------------------------------------------

def _expand(desc):
    names = []
    for s in desc["species"]:
        names += [s["name"]] * s["count"]
    return names


def _golden_names(case, desc):
    """Ion names in index order, and the species list, from the reference's files."""
    d = os.path.join(INI, case)
    files = []
    for root, _, fs in os.walk(d):
        files += [os.path.join(root, f) for f in fs if not f.endswith("_logfile.txt")]
    names = {}
    for path in files:
        if "_cloud" in os.path.basename(path):
            continue
        for line in open(path):
            if line[:1].isdigit():
                k, name = line.split()[:2]
                names[int(k)] = name
    species = {}
    for path in files:
        if "_cloud" not in os.path.basename(path):
            continue
        for line in open(path):
            if line[:1].isdigit():
                c = line.split()
                species[int(c[0])] = (c[1], int(c[2]))
    return [names[k] for k in sorted(names)], [species[k] for k in sorted(species)]


@pytest.mark.parametrize("case", ["paul_cycle", "adaptive_separate", "two_ramps", "adaptive_ramp"])
def test_composition_matches_reference_program(case):
    desc = S.describe_ini(os.path.join(INI, case + ".ini"))
    names, species = _golden_names(case, desc)
    assert _expand(desc) == names
    assert [s["name"] for s in desc["species"]] == [s[0] for s in species]
    assert all(n == desc["N"] for _, n in species)


def _oracle_run(case, ramp_end_at_lifetime=False):
    """The runner's operation loop restated over the oracle's own step()
    (bin/ode.cpp:459-494): prints after the step that passes each print time,
    after every operation, and at the end; a ramp leaves Vrf/Vdc at the values
    of its last step's start.  With ramp_end_at_lifetime the ramp ends at the
    run's lifetime instead: a planted mistake that the golden files must
    reject.  Returns [(lifetime, pos, vel)] per print."""
    d = S.describe_ini(os.path.join(INI, case + ".ini"))
    amu = np.array([sp["mass_amu"] for sp in d["species"] for _ in range(sp["count"])])
    ch = np.array([sp["charge"] for sp in d["species"] for _ in range(sp["count"])], dtype=np.int32)
    pos, vel = S.init_cloud(d["N"], amu, (d["temp"],) * 3, d["semiaxis"], d["offset"], seed=0)
    tr = d["trap"]
    base = dict(kappa=tr["kappa"], r0=tr["r0"], z0=tr["z0"], freq_rf=tr["freq_rf"],
                E_kick=tr["E_kick"], newVrf=tr["newVrf"], newVdc=tr["newVdc"],
                coulomb=1 if d["coulomb"] else 0, coulomb_scale=d["coulomb_scale"])
    c = O.Cloud(pos, vel, O.mass_kg(amu), ch, O.fields(Vrf=tr["Vrf"], Vdc=tr["Vdc"], **base),
                h0=d["timestep"], adaptive=d["adaptive"], time_mode=O.TIME_FROZEN)
    vrf, vdc, interval, nr, t = tr["Vrf"], tr["Vdc"], d["print_interval"], 0, 0.0
    prints = []
    for op in d["operations"]:
        assert op["ode_timestep"] == 0 and op["type"] in ("normal", "trap_ramp", "tof", "free")
        if op["print_timestep"]:
            interval = op["print_timestep"]
        t0, end = t, t + op["duration"]
        ramp = op["type"] == "trap_ramp"
        kind = dict(trap=1 if op["type"] in ("normal", "trap_ramp") else 0,
                    tof=1 if op["type"] == "tof" else 0)
        c.set_fields(O.fields(Vrf=vrf, Vdc=vdc, **kind, **base))
        step_vrf, step_vdc = vrf, vdc
        while t < end:
            if ramp:   # the trap's values at this step's start (bin/ion_fly.cpp:95-99)
                x = (t - t0) / (end - t0)
                step_vrf = vrf * (1 - x) + tr["newVrf"] * x
                step_vdc = vdc * (1 - x) + tr["newVdc"] * x
                c.set_fields(O.fields(Vrf=step_vrf, Vdc=step_vdc, **kind, **base))
            c.step()
            p_, v_, sc = c.state()
            t = sc["lifetime"]
            if t > interval * nr:
                prints.append((t, p_, v_))
                nr += 1
        if ramp:
            if ramp_end_at_lifetime:
                x = (t - t0) / (end - t0)
                step_vrf = vrf * (1 - x) + tr["newVrf"] * x
                step_vdc = vdc * (1 - x) + tr["newVdc"] * x
            vrf, vdc = step_vrf, step_vdc
        if d["print_after_operation"]:
            p_, v_, _ = c.state()
            prints.append((t, p_, v_))
    p_, v_, _ = c.state()
    prints.append((t, p_, v_))
    return prints


def _golden_particle_rows(case):
    """[(t, pos, vel)] per print from the reference's files (ms and mm back to s and m)."""
    rows = {}
    for root, _, fs in os.walk(os.path.join(INI, case)):
        for f in fs:
            if "_cloud" in f or f.endswith("_logfile.txt"):
                continue
            for ln in open(os.path.join(root, f)):
                if ln[:1].isdigit():
                    w = ln.split()
                    rows.setdefault(int(w[0]), []).append([float(x) for x in w[2:9]])
    a = np.array([rows[k] for k in sorted(rows)])          # [ion, print, 7]
    return [(a[0, j, 0] * 1e-3, a[:, j, 1:4] * 1e-3, a[:, j, 4:7]) for j in range(a.shape[1])]


def _prints_agree(mine, ref, rtol=3e-6):
    if len(mine) != len(ref):
        return False
    for (tm, pm, vm), (tr_, pr, vr) in zip(mine, ref):
        if abs(tm - tr_) > rtol * tr_:
            return False
        for a, b in ((pm, pr), (vm, vr)):
            if np.any(np.abs(a - b) > rtol * np.abs(b) + rtol * np.abs(b).max()):
                return False
    return True


@pytest.mark.parametrize("case", ["paul_cycle", "two_ramps", "adaptive_ramp"])
def test_operation_rules_over_the_oracle(case):
    """The print schedule and the ramp rule, restated over the oracle, give
    the reference program's own files; ending the ramp at the lifetime
    instead of the last step's start does not (where a ramp runs)."""
    ref = _golden_particle_rows(case)
    assert _prints_agree(_oracle_run(case), ref)
    if case != "paul_cycle":
        assert not _prints_agree(_oracle_run(case, ramp_end_at_lifetime=True), ref)


def test_describe_golden_files():
    d = S.describe_ini(os.path.join(INI, "paul_cycle.ini"))
    assert d["N"] == 12 and d["coulomb"] and d["coulomb_scale"] == 1.0
    assert d["trap"]["newVrf"] == 80 and d["trap"]["E_kick"] == 100
    assert [o["type"] for o in d["operations"]] == ["normal", "trap_ramp", "tof", "free"]
    d = S.describe_ini(os.path.join(INI, "adaptive_separate.ini"))
    assert d["semiaxis"][0] == 4e-4                                 # "sx: ..." pair
    assert [(s["name"], s["count"], s["charge"]) for s in d["species"]] == \
        [("40Zz2+", 3, 2), ("88Ww+", 6, 1)]
    assert d["adaptive"] and d["separate"] and d["coulomb_scale"] == 2.0
    # operation5's unknown type is skipped; "trap" and "none" are normal operations
    assert [o["type"] for o in d["operations"]] == ["normal", "tof", "normal"]
    assert d["operations"][0]["ode_timestep"] == 1.0
    d = S.describe_ini(os.path.join(INI, "two_ramps.ini"))
    assert d["N"] == 7 and d["temp"] == 0.005 and d["trap"]["Vrf"] == 40   # keys ignore case
    assert not d["coulomb"] and not d["adaptive"]
    # the last ramp's targets serve both ramps
    assert (d["trap"]["newVrf"], d["trap"]["newVdc"]) == (20.0, 6.0)
    assert [s["count"] for s in d["species"]] == [2, 2, 3]


BASE = """[cloud]
NParticles = 10
[composition]
133Xx+ = 1
[ode]
method = DP
[output]
prefix = p
[operation0]
type = normal
"""


def _describe(tmp_path, text, name="s.ini"):
    p = tmp_path / name
    p.write_bytes(text.encode() if isinstance(text, str) else text)
    return S.describe_ini(str(p))


def test_ini_defaults(tmp_path):
    d = _describe(tmp_path, BASE)
    # DoSimulation's defaults (bin/initialization.cpp:310-317,428-505,514-516)
    assert d["temp"] == 1e-3 and d["semiaxis"] == [1e-3] * 3 and d["offset"] == [0.0] * 3
    assert d["timestep"] == 1e-9 and d["adaptive"] is False
    assert d["coulomb"] is False and d["coulomb_scale"] == 0.0
    assert d["print_interval"] == 1e-6 and not d["separate"] and not d["print_after_operation"]
    assert d["operations"] == [{"type": "normal", "duration": 1e-3, "print_timestep": 0.0,
                                "ode_timestep": 0.0}]
    # no [trap] section: the trap parameters' own defaults (bin/trapparameters.cpp:2-10)
    f = S.default_fields()
    assert d["trap"]["kappa"] == f.kappa and d["trap"]["freq_rf"] == f.freq_rf
    # a [trap] section brings the reader's defaults (bin/initialization.cpp:449-454)
    d = _describe(tmp_path, BASE + "[trap]\ntype = ideal\n")
    assert (d["trap"]["kappa"], d["trap"]["freq_rf"], d["trap"]["Vrf"]) == (0.025, 0.04, 50.0)


def test_ini_syntax(tmp_path):
    text = (
        "﻿; a comment\n"
        "# another\n"
        "[cloud]\n"
        "NParticles = 10\n"
        "  temp = 0.5\n"                      # indented pair: a continuation of NParticles
        "sx : 0.002 ; inline comment\n"
        "sy = 0.003;not a comment\n"          # ';' without a blank before it is text
        "sz = 0.004\n"
        "sz = 0.009\n"                        # repeated name: joined with '\\n', strtod reads the first
        "this line has no separator\n"
        "[composition]\n"
        "133Xx+ = 1\n"
        "[ode]\nmethod = DP\n[output]\nprefix = p\n[operation0]\ntype = normal\n")
    d = _describe(tmp_path, text)
    assert d["N"] == 10                      # "10\ntemp = 0.5" reads as 10
    assert d["temp"] == 1e-3                 # the continuation set no temp
    assert d["semiaxis"] == [0.002, 0.003, 0.004]
    assert d["skipped_line"] == 10


INI_EDGE_CASES = [
    "\ufeff[cloud]\nNParticles = 12\n",
    "; comment\n# comment\n   ; indented comment\n[cloud]\nsx = 1 ; inline\nsy = 2;not inline\n"
    "sz : 3\nx0:4\n",
    "[cloud]\nNParticles = 10\n  temp = 0.5\n\tmore\n[ode]\n  method = DP\n",
    "[cloud]\nsx = 1\nsx = 2\nSX = 3\nNParticles = 4\nnparticles = 5\n",
    "[Cloud]\nsx = 1\n[cloud]\nsy = 2\n[CLOUD]\nsx = 9\n",
    "[cloud]\nsx = 0.002" + " " * 189 + "sy = 0.007\n",
    "[cloud]\nsx = 0.002" + " " * 195 + "sy = 0.007\n",
    "[cloud]\n" + "k" * 60 + " = 1\n  continued\n",
    "[" + "s" * 60 + "]\nx = 1\n",
    "[cloud]\r\nsx = 1\r\nsy = 2 \r\n",
    "[cloud]\n= nameless\nempty =\na = b = c\n   \n\t\n",
    "[cloud]\nno separator\n[unclosed\nsx = 5\n[ok]\nx=1\n",
    "x = before any section\n[cloud]\nsx = 1\n",
    "[cloud]\nsx = 1\n[cloud ; comment]\nsy = 2\n[cloud]x = 3\n",
]
INI_QUERIES = [("cloud", "NParticles"), ("cloud", "temp"), ("cloud", "sx"), ("cloud", "sy"),
               ("cloud", "sz"), ("cloud", "x0"), ("Cloud", "SX"), ("ode", "method"),
               ("cloud", "k" * 60), ("cloud", "k" * 49), ("s" * 49, "x"), ("s" * 60, "x"),
               ("cloud", ""), ("cloud", "empty"), ("cloud", "a"), ("ok", "x"), ("", "x"),
               ("cloud", "no separator"), ("cloud ; comment", "sy"), ("cloud", "x"),
               ("cloud ", "sy")]


@have_refsim
@pytest.mark.parametrize("k", range(len(INI_EDGE_CASES)))
def test_ini_reader_matches_reference(tmp_path, k):
    """The reader against the reference's own INIReader (bin/ini.c and
    bin/inireader.cpp, compiled into oracle/_ref/ref_init) on edge cases of
    the format: same sections, same field sets, same Get() values."""
    path = tmp_path / "edge.ini"
    path.write_bytes(INI_EDGE_CASES[k].encode("utf-8"))
    ref_sections, ref = O.ref_ini(path, INI_QUERIES)
    assert S.ini_names(path) == ref_sections
    for (sec, name), (fields, value) in zip(INI_QUERIES, ref):
        assert S.ini_names(path, sec) == fields, (sec, name)
        assert S.ini_value(path, sec, name) == value, (sec, name)


def test_ini_long_lines_and_sections(tmp_path):
    # inih reads 199-character chunks: the rest of a long line is a line of its own ...
    long = "sx = 0.002" + " " * 189 + "sy = 0.007\n"
    d = _describe(tmp_path, BASE.replace("NParticles = 10\n", "NParticles = 10\n" + long))
    assert d["semiaxis"][:2] == [0.002, 0.007]
    # ... and a continuation of the previous name when it starts with blanks
    long = "sx = 0.002" + " " * 195 + "sy = 0.007\n"
    d = _describe(tmp_path, BASE.replace("NParticles = 10\n", "NParticles = 10\n" + long))
    assert d["semiaxis"][:2] == [0.002, 1e-3]
    # section names are case-sensitive: [Cloud] is not [cloud]
    with pytest.raises(S.SimbucaError, match=r"no \[cloud\]"):
        _describe(tmp_path, BASE.replace("[cloud]", "[Cloud]"))


def test_ini_trap_types(tmp_path):
    # only "ideal" reads [trap]; any other type keeps the trap's defaults
    d = _describe(tmp_path, BASE + "[trap]\ntype = real\nVrf = 99\n")
    assert d["trap"]["Vrf"] == S.default_fields().Vrf and d["trap"]["kappa"] == S.default_fields().kappa
    d = _describe(tmp_path, BASE + "[trap]\nVrf = 99\n")
    assert d["trap"]["Vrf"] == 99.0


def test_ini_composition_rules(tmp_path):
    comp = "[composition]\n133Xx+ = 0.25\n85Yy+ = 0.25\n39Zz2+ = 0.9\n"
    d = _describe(tmp_path, BASE.replace("[composition]\n133Xx+ = 1\n", comp))
    # sorted names; n = (int)(fraction N), capped by what is left; the last takes the rest
    assert [(s["name"], s["count"]) for s in d["species"]] == \
        [("133Xx+", 2), ("39Zz2+", 8)]
    with pytest.raises(S.SimbucaError, match="fraction"):
        _describe(tmp_path, BASE.replace("133Xx+ = 1", "133Xx+ = 1.5"))
    # a negative fraction takes every ion left (an unsigned comparison in the reference)
    comp = "[composition]\n133Xx+ = -0.1\n85Yy+ = 0.5\n"
    d = _describe(tmp_path, BASE.replace("[composition]\n133Xx+ = 1\n", comp))
    assert [(s["name"], s["count"]) for s in d["species"]] == [("133Xx+", 10)]
    with pytest.raises(S.SimbucaError, match="NParticles"):
        _describe(tmp_path, BASE.replace("NParticles = 10", "NParticles = 1e12"))


def test_ini_operations(tmp_path):
    ops = ("[operation0]\ntype = tof\nEkick = 5\n"
           "[operation3]\ntype = tof\nEkick = 7\nduration = 2e-6\n"
           "[operation4]\ntype = trap_ramp\nVrf = 1\nVdc = 2\n"
           "[operation9]\ntype = nop\nprint_timestep = 3e-7\n"
           "[operation31]\ntype = none\n"
           "[operation32]\ntype = normal\n")           # beyond MAX_OPERATIONS: not read
    d = _describe(tmp_path, BASE.replace("[operation0]\ntype = normal\n", ops))
    assert [o["type"] for o in d["operations"]] == ["tof", "tof", "trap_ramp", "free", "normal"]
    assert d["trap"]["E_kick"] == 7.0                  # the last Ekick serves every tof
    assert d["operations"][1]["duration"] == 2e-6
    assert d["operations"][3]["print_timestep"] == 3e-7


@pytest.mark.parametrize("edit,msg", [
    (("[ode]\nmethod = DP\n", ""), "ODE configuration"),
    (("method = DP", "method = RK"), "only Dormand-Prince"),
    (("[output]\nprefix = p\n", ""), "Output file"),
    (("type = normal", "type = whatever"), "No operation"),
    (("[composition]\n133Xx+ = 1\n", ""), "composition"),
    (("type = normal", "type = dissociation\nproduct = two words"), "one word"),
    (("[output]\n", "[import]\nprefix = old\n[output]\n"), "cannot open old_pAll"),
    (("prefix = p", "prefix = p\ndigits = 18"), "digits"),
    (("type = normal", "type = normal\node_timestep = 1e-9"), "print_timestep"),
])
def test_ini_not_provided_or_wrong(tmp_path, edit, msg):
    with pytest.raises(S.SimbucaError, match=msg):
        _describe(tmp_path, BASE.replace(*edit))


def test_ini_import_reads_a_print(tmp_path, rng):
    """[import]: the cloud is the last print of <prefix>_pAll.txt (the writer's
    11 columns), names through the composition rules, lifetime from its t."""
    n = 5
    names = ["133Xx+", "133Xx+", "40Zz2+", "7Ww-", "133Xx+"]
    amu = np.array([133.0, 133.0, 40.0, 7.0, 133.0])
    pos = rng.normal(0, 1e-3, (n, 3)); vel = rng.normal(0, 100.0, (n, 3))
    w = S.Writer(str(tmp_path / "snap"), names, amu, digits=17)
    w.particles(1e-6, pos * 0.5, vel * 0.5)
    w.close(2.5e-6, pos, vel)
    text = BASE.replace("[cloud]\nNParticles = 10\n[composition]\n133Xx+ = 1\n",
                        "[import]\nprefix = %s\n" % (tmp_path / "snap"))
    d = _describe(tmp_path, text)
    assert d["import"] == str(tmp_path / "snap") and d["N"] == n
    assert d["lifetime"] == pytest.approx(2.5e-6, rel=1e-15)
    assert [(s["name"], s["count"], s["mass_amu"], s["charge"]) for s in d["species"]] == [
        ("133Xx+", 2, 133.0, 1), ("40Zz2+", 1, 40.0, 2), ("7Ww-", 1, 7.0, -1), ("133Xx+", 1, 133.0, 1)]
    # a name with no mass is an error, as the reference's importer stops there
    w = S.Writer(str(tmp_path / "bad"), ["Qq+"], [1.0])
    w.close(1e-6, np.zeros((1, 3)), np.zeros((1, 3)))
    with pytest.raises(S.SimbucaError, match="no mass"):
        _describe(tmp_path, text.replace("snap", "bad"))


def test_ini_dissociation_and_print_at_x(tmp_path):
    ops = ("[operation0]\ntype = dissociation\nreactant = 133Xx+\nproduct = 85Pd+\n"
           "fraction = 0.4\nmass = 1.4e-25\n"
           "[operation1]\ntype = normal\n"
           "[operation2]\ntype = dissociation\n")
    text = BASE.replace("[operation0]\ntype = normal\n", ops).replace(
        "[output]\n", "[output]\nprint_at_x = yes\nprint_x = 2e-4\n")
    d = _describe(tmp_path, text)
    assert d["print_at_x"] is True and d["print_x"] == 2e-4
    assert [o["type"] for o in d["operations"]] == ["dissociation", "normal", "dissociation"]
    o = d["operations"][0]
    assert (o["fraction"], o["reactant"], o["product"], o["mass"]) == (0.4, "133Xx+", "85Pd+", 1.4e-25)
    # the reader's defaults (bin/initialization.cpp:530-534)
    o = d["operations"][2]
    assert (o["fraction"], o["reactant"], o["product"], o["mass"]) == (1.0, "none", "none", 0.0)
    d = _describe(tmp_path, BASE)
    assert d["print_at_x"] is False and d["print_x"] == 0.0


def test_writer_rename_and_x_prints(tmp_path):
    w = S.Writer(str(tmp_path / "w"), ["133Xx+", "85Yy+"], [133.0, 85.0])
    pos = np.array([[1e-3, 2e-3, 3e-3], [-1e-3, 0.0, 1e-3]]); vel = np.array([[10.0, 0, 0], [0, 5.0, 0]])
    w.x(0, 1e-6, pos, vel)                        # opens the file: header only
    w.x(1, 2e-6, pos, vel)
    w.set_ion(0, "40Pd+", 40 * S.AMU)
    w.particles(3e-6, pos, vel)
    w.cloud(3e-6, pos, vel)
    w.close()
    xl = [ln.split() for ln in open(tmp_path / "w_x.txt") if ln[:1].isdigit()]
    assert xl == [["1", "85Yy+", "2.000000e-03", "-1.000000e+00", "0.000000e+00", "1.000000e+00",
                   "0.000000e+00", "5.000000e+00", "0.000000e+00"]]
    assert open(tmp_path / "w_x.txt").readline().startswith("#Print at X position#")
    pl = [ln.split() for ln in open(tmp_path / "w_pAll.txt") if ln[:1].isdigit()]
    assert pl[0][1] == "40Pd+" and float(pl[0][-1]) == pytest.approx(0.5 * 40 * S.AMU * 100 / S.EL_CHARGE,
                                                                      rel=1e-5)
    # species files keep the opening species: the renamed ion leaves the 133Xx+ sums
    c0 = [ln.split() for ln in open(tmp_path / "w_cloud0.txt") if ln[:1].isdigit()]
    assert c0[0][1] == "133Xx+" and float(c0[0][4]) == 0.0 and float(c0[0][-1]) == 0.0
    with pytest.raises(S.SimbucaError, match="one non-empty word"):
        S.Writer(str(tmp_path / "v"), ["a"], [1.0]).set_ion(0, "two words", 1.0)


EXE = os.path.join(os.path.dirname(HERE), "simbuca_amd", "simbuca_hip")


@pytest.mark.skipif(not os.path.exists(EXE), reason="build it with make -C simbuca_amd/csrc")
def test_program_describe(tmp_path):
    """simbuca_hip --describe (no GPU): the settings as JSON; the mass tables
    come from ../libraries as the reference's, when present."""
    p = subprocess.run([EXE, "--describe", os.path.join(INI, "two_ramps.ini")], capture_output=True,
                       text=True, timeout=60)
    assert p.returncode == 0, p.stderr
    assert json.loads(p.stdout) == S.describe_ini(os.path.join(INI, "two_ramps.ini"))
    p = subprocess.run([EXE, "--describe", str(tmp_path / "none.ini")], capture_output=True,
                       text=True, timeout=60)
    assert p.returncode == 1 and "cannot read" in p.stderr
    if os.path.exists(AME) and os.path.exists(PDG):
        run = tmp_path / "run"
        run.mkdir()
        os.symlink(O.REF_LIBRARIES, tmp_path / "libraries")
        (run / "s.ini").write_text(BASE.replace("133Xx+", "133Cs+"))
        p = subprocess.run([EXE, "--describe", "s.ini"], cwd=str(run), capture_output=True,
                           text=True, timeout=60)
        assert p.returncode == 0, p.stderr
        t = S.MassTable(AME, PDG)
        assert json.loads(p.stdout)["species"][0]["mass_amu"] == t.mass("133Cs")


def test_ini_missing_file(tmp_path):
    with pytest.raises(S.SimbucaError, match="cannot read"):
        S.describe_ini(str(tmp_path / "none.ini"))
