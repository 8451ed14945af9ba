"""The C ABI driven from C++ (tests/cxx/abi_caller.cpp, built with g++
against include/simbuca_hip.h only), the way the reference would link it:
the CUNBODY begin/end pattern of bin/force.cpp:48-93 and the engine's
sbk_step, both checked against the CPU oracle."""
import os
import subprocess
import tempfile

import numpy as np
import pytest

import oracle as O
from conftest import ROOT

pytestmark = pytest.mark.gpu

CALLER = os.path.join(ROOT, "tests", "cxx", "abi_caller")


def _run(mode, payload, nout):
    assert os.path.exists(CALLER), "build tests/cxx first (make -C tests/cxx)"
    with tempfile.TemporaryDirectory() as d:
        fi, fo = os.path.join(d, "in.bin"), os.path.join(d, "out.bin")
        with open(fi, "wb") as fh:
            fh.write(payload)
        p = subprocess.run([CALLER, mode, fi, fo], capture_output=True, text=True, timeout=120)
        assert p.returncode == 0, p.stderr
        out = np.fromfile(fo, dtype=np.float64)
    assert out.size == nout
    return out


def test_cunbody_caller_matches_cpu_force(rng):
    n = 777
    amu = np.full(n, 133.0)
    ch = np.ones(n, dtype=np.int32)          # the CUNBODY path ignores charges (bin/force.cpp:89)
    pos = rng.normal(0, 5e-4, (n, 3)); vel = rng.normal(0, 137.0, (n, 3))
    f = O.fields(trap=1, Vrf=50.0, Vdc=5.0, kappa=0.001, r0=0.04, z0=0.04, freq_rf=5e6, tof=1,
                 E_kick=3.0, coulomb=1, coulomb_scale=1.0)
    out = _run("cunbody", O._ref_input(1, amu, ch, pos, vel, f, 1.1e-6), 3 * n).reshape(n, 3)
    ref = O.force(pos, vel, O.mass_kg(amu), ch, f, 1.1e-6)
    # same physics, different association of the prefactor and a 1e-23 m^2 softening
    assert np.allclose(out, ref, rtol=1e-12, atol=1e-12 * np.abs(ref).max())


def test_engine_caller_matches_oracle_steps(rng):
    n = 96
    amu = np.full(n, 133.0)
    ch = np.ones(n, dtype=np.int32)
    pos = rng.normal(0, 5e-4, (n, 3)); vel = rng.normal(0, 137.0, (n, 3))
    f = O.fields(trap=1, Vrf=0.0, Vdc=5.0, kappa=0.001, r0=0.04, z0=0.04, coulomb=1)
    nsteps = 60
    payload = O._ref_input(2, amu, ch, pos, vel, f, 0.0, h0=1e-9, rtol=1e-7, atol=1e-7,
                           adaptive=True, nsteps=nsteps)
    out = _run("engine", payload, 6 * n + 2)
    c = O.Cloud(pos, vel, O.mass_kg(amu), ch, f, h0=1e-9, rtol=1e-7, atol=1e-7)
    for _ in range(nsteps):
        c.step()
    op, ov, sc = c.state()
    assert out[6 * n + 1] == sc["attempts"]
    assert out[6 * n] == pytest.approx(sc["lifetime"], rel=1e-10)
    assert np.allclose(out[:3 * n].reshape(n, 3), op, rtol=1e-9, atol=1e-9 * np.abs(op).max())
    assert np.allclose(out[3 * n:6 * n].reshape(n, 3), ov, rtol=1e-9, atol=1e-9 * np.abs(ov).max())
