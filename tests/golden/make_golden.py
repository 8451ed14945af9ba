"""Writes the golden fixtures of tests/golden/ by running the reference's own
code, compiled from its sources by `make -C oracle ref refsim`:

  force, steps  its CPU force()/step() (oracle/_ref/ref_harness);
  output        its IonCloud writer (output_n7/);
  init          its InitCloud (init_cloud.npz) and mass tables (masses.json),
                through oracle/_ref/ref_init;
  ini           the reference program itself (oracle/_ref/simbuca) on the
                simulation files ini/<case>.ini: the text files it writes go
                to ini/<case>/ (its _bundled.txt summary, which reads columns
                its writer does not write, is left out).

Each fixture holds the inputs and the reference's outputs as plain arrays
(numpy .npz, no pickles), JSON or the reference's own text files.  Run from
the repository root:

    make -C oracle ref refsim && python tests/golden/make_golden.py [force] [steps] [output] [init] [ini]
"""
from __future__ import annotations

import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402

OUT = os.path.dirname(os.path.abspath(__file__))

TRAP_KEYS = ["trap", "kappa", "Vdc", "Vrf", "r0", "z0", "freq_rf", "tof", "E_kick", "coulomb",
             "coulomb_scale"]


def _fields_arr(f):
    return np.array([float(getattr(f, k)) for k in TRAP_KEYS])


def force_case(name, seed, n, sigma, species, charges, f, lifetime):
    rng = np.random.Generator(np.random.PCG64(seed))
    amu = rng.choice(np.array(species, dtype=np.float64), n)
    ch = rng.choice(np.array(charges, dtype=np.int32), n).astype(np.int32)
    pos = rng.normal(0.0, sigma, (n, 3))
    vel = rng.normal(0.0, 137.0, (n, 3))
    out = oracle.ref_force(amu, ch, pos, vel, f, lifetime)
    np.savez(os.path.join(OUT, name), mass_amu=amu, charge=ch, pos=pos, vel=vel,
             fields=_fields_arr(f), lifetime=np.float64(lifetime), derivs=out)
    print(name, n, float(np.abs(out).max()))


def steps_case(name, seed, n, sigma, f, h0, rtol, atol, adaptive, nsteps, lifetime=0.0):
    rng = np.random.Generator(np.random.PCG64(seed))
    amu = np.full(n, 133.0)
    ch = np.ones(n, dtype=np.int32)
    pos = rng.normal(0.0, sigma, (n, 3))
    vel = rng.normal(0.0, 137.0, (n, 3))
    p, v, sc, hist = oracle.ref_steps(amu, ch, pos, vel, f, lifetime, h0, rtol, atol, adaptive,
                                      nsteps)
    np.savez(os.path.join(OUT, name), mass_amu=amu, charge=ch, pos=pos, vel=vel,
             fields=_fields_arr(f), lifetime=np.float64(lifetime), h0=np.float64(h0),
             rtol=np.float64(rtol), atol=np.float64(atol), adaptive=np.int32(adaptive),
             nsteps=np.int32(nsteps), pos_out=p, vel_out=v,
             scalars=np.array([sc["lifetime"], sc["h"], sc["hnext"], sc["errold"], sc["reject"],
                               sc["countstep"], sc["err"]]),
             hist=hist)
    print(name, n, nsteps, sc)


def output_case(name, seed, n, species, prints, lifetime):
    """The reference's own output files (IonCloud::PrintParticles, PrintCloud,
    Delete) for a small cloud, in both file layouts, with the inputs."""
    import shutil
    import tempfile
    rng = np.random.Generator(np.random.PCG64(seed))
    amu = rng.choice(np.array(species, dtype=np.float64), n)
    ch = rng.choice(np.array([1, 1, 2], dtype=np.int32), n).astype(np.int32)
    pos = rng.normal(0.0, 1e-3, (n, 3))
    vel = rng.normal(0.0, 137.0, (n, 3))
    base = os.path.join(OUT, name)
    np.savez(base + ".npz", mass_amu=amu, charge=ch, pos=pos, vel=vel,
             lifetime=np.float64(lifetime), prints=np.int32(prints))
    for layout, per_particle in (("pAll", False), ("per_particle", True)):
        d = os.path.join(base, layout)
        shutil.rmtree(d, ignore_errors=True)
        os.makedirs(d)
        with tempfile.TemporaryDirectory() as tmp:
            files = oracle.ref_output(amu, ch, pos, vel, lifetime, prints, tmp,
                                      per_particle=per_particle)
        for fname, text in files.items():
            with open(os.path.join(d, "ref" + fname), "w") as fh:
                fh.write(text)


INIT_CASES = [
    # (seed, energies, semiaxis, offset): one per InitCloud branch
    (0, (0.1,), (1e-3, 1e-3, 1e-3), (0.0, 0.0, 0.0)),
    (12345, (0.05, 0.2), (4e-4, 2e-3, 7e-4), (1e-4, -2e-4, 3e-4)),
    (77, (0.01, 0.01, 0.01), (2e-3, 5e-4, 1e-3), (-1e-3, 0.0, 2e-4)),
]
MASS_NAMES = ["133Cs", "85Rb", "87Rb", "39K", "1H", "2H", "3H", "3He", "4He", "7Li", "12C", "16O",
              "23Na", "40Ar", "40Ca", "88Sr", "138Ba", "208Pb", "238U", "1n", "proton", "neutron",
              "e-", "e+", "mu-", "pi+", "133Xx", "Cs"]
INI_CASES = ["paul_cycle", "adaptive_separate", "two_ramps", "tof_detector", "adaptive_ramp"]


def init_case(name="init_cloud.npz", n=240):
    rng = np.random.Generator(np.random.PCG64(21))
    amu = rng.choice(np.array([133.0, 85.0, 39.0, 7.016, 1.00727]), n)
    arrays = {"mass_amu": amu}
    for k, (seed, e, semi, off) in enumerate(INIT_CASES):
        pos, vel = oracle.ref_init_cloud(n, seed, e, semi, off, amu)
        arrays.update({f"seed{k}": np.int64(seed), f"energy{k}": np.array(e),
                       f"semiaxis{k}": np.array(semi), f"offset{k}": np.array(off),
                       f"pos{k}": pos, f"vel{k}": vel})
    np.savez(os.path.join(OUT, name), **arrays)
    print(name, n, len(INIT_CASES))


def masses_case(name="masses.json"):
    import json
    got = oracle.ref_masses(MASS_NAMES)
    rows = [{"name": k, "table": got[k][0],
             "mass_amu": None if got[k][0] == "none" else repr(got[k][1])} for k in MASS_NAMES]
    with open(os.path.join(OUT, name), "w") as fh:
        json.dump(rows, fh, indent=1)
        fh.write("\n")
    print(name, len(rows))


def ini_case(case):
    import shutil
    with open(os.path.join(OUT, "ini", case + ".ini")) as fh:
        text = fh.read()
    files = oracle.ref_run_ini(text)
    d = os.path.join(OUT, "ini", case)
    shutil.rmtree(d, ignore_errors=True)
    kept = 0
    for rel, body in sorted(files.items()):
        if rel.endswith("_bundled.txt"):
            continue
        path = os.path.join(d, rel)
        os.makedirs(os.path.dirname(path), exist_ok=True)
        with open(path, "w") as fh:
            fh.write(body)
        kept += 1
    print("ini", case, kept, "files")


def main(which=("force", "steps", "output")):
    if "init" in which or "ini" in which:
        if not oracle.refsim_available():
            raise SystemExit("oracle/_ref/simbuca or ref_init is not built (make -C oracle refsim)")
        if "init" in which:
            init_case()
            masses_case()
        if "ini" in which:
            for case in INI_CASES:
                ini_case(case)
    if not oracle.ref_available():
        raise SystemExit("oracle/_ref/ref_harness is not built (make -C oracle ref)")
    if "output" in which:
        output_case("output_n7", 15, 7, [133.0, 85.0, 39.0], 2, 2.5e-6)
    if "force" not in which and "steps" not in which:
        return
    # force(): Paul-trap RF+DC at t != 0, TOF kick, Coulomb; mixed species and charges
    f = oracle.fields(trap=1, Vrf=50.0, Vdc=5.0, kappa=0.001, r0=0.04, z0=0.04, freq_rf=50e5,
                      tof=1, E_kick=3.0, coulomb=1, coulomb_scale=1.0)
    force_case("force_n500_trap_tof_coulomb.npz", 11, 500, 5e-4, [133.0, 85.0], [1, 1, 2, -1],
               f, 1.234e-6)
    # force(): Coulomb only, scaled, ragged size
    f = oracle.fields(trap=0, coulomb=1, coulomb_scale=0.5)
    force_case("force_n257_coulomb_scaled.npz", 12, 257, 1e-3, [39.0, 133.0], [1], f, 0.0)
    # 200 adaptive steps at the reference tolerances (rtol = atol = 1e-7)
    f = oracle.fields(trap=1, Vrf=0.0, Vdc=5.0, kappa=0.001, r0=0.04, z0=0.04, coulomb=1)
    steps_case("steps_adaptive_n64.npz", 13, 64, 5e-4, f, 1e-9, 1e-7, 1e-7, True, 200)
    # 20 fixed steps with a time-dependent RF field frozen at t_n in every stage
    f = oracle.fields(trap=1, Vrf=50.0, Vdc=5.0, kappa=0.001, r0=0.04, z0=0.04, freq_rf=50e5,
                      coulomb=1)
    steps_case("steps_fixed_rf_n64.npz", 14, 64, 5e-4, f, 2e-9, 1e-7, 1e-7, False, 20,
               lifetime=3.0e-7)


if __name__ == "__main__":
    main(tuple(sys.argv[1:]) or ("force", "steps", "output"))
