"""CPU-side checks of the product library: it builds, loads, and exports every
symbol include/simbuca_hip.h declares; host-side helpers agree with the
oracle.  No GPU work is launched here."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

import oracle as O
import simbuca_amd as S
from simbuca_amd import synth
from conftest import ROOT


def _declared_symbols(header="simbuca_hip.h"):
    src = open(os.path.join(ROOT, "include", header)).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    names = re.findall(r"^\s*(?:void|int|float|const char)\s*\*?\s*(\w+)\s*\(", src, flags=re.M)
    return sorted(set(names))


def test_header_and_binding_agree():
    assert _declared_symbols() == sorted(S.EXPORTS)


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(S.LIB_PATH)
    for name in _declared_symbols():
        assert hasattr(lib, name), name


def test_nbody_interface_is_a_library_of_its_own():
    # the reference's NBODY names (bin/nbody.cpp:12-18) are generic, so they
    # live in libsimbuca_nbody.so and the engine library exports none of them
    assert _declared_symbols("simbuca_nbody.h") == sorted(S.NBODY_EXPORTS)
    nb = ctypes.CDLL(S.NBODY_LIB_PATH)
    for name in S.NBODY_EXPORTS:
        assert hasattr(nb, name), name
    out = subprocess.run(["nm", "-D", "--defined-only", S.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout.split()
    for name in S.NBODY_EXPORTS:
        assert name not in out, name


def _c_layout(tmp_path):
    """sizeof and offsetof of the header's structs, as gcc lays them out."""
    src = tmp_path / "layout.c"
    exe = tmp_path / "layout"
    fields = {"sbk_config": [n for n, _ in S.Config._fields_],
              "sbk_cloud_status": [n for n, _ in S.CloudStatus._fields_],
              "sbk_stats": [n for n, _ in S.Stats._fields_],
              "sbk_fields": [n for n, _ in S.Fields._fields_]}
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "simbuca_hip.h"', 'int main(void) {']
    for t, names in fields.items():
        lines.append(f'printf("{t} size %zu\\n", sizeof({t}));')
        for n in names:
            lines.append(f'printf("{t} {n} %zu\\n", offsetof({t}, {n}));')
    lines.append("return 0; }")
    src.write_text("\n".join(lines))
    subprocess.run(["gcc", "-std=c99", "-I", os.path.join(ROOT, "include"), "-o", str(exe), str(src)],
                   check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout
    return {tuple(ln.split()[:2]): int(ln.split()[2]) for ln in out.splitlines()}


def test_struct_layouts_match_header(tmp_path):
    # every ctypes mirror has the C struct's size and field offsets
    lay = _c_layout(tmp_path)
    for t, cls in (("sbk_config", S.Config), ("sbk_cloud_status", S.CloudStatus),
                   ("sbk_stats", S.Stats), ("sbk_fields", S.Fields)):
        assert ctypes.sizeof(cls) == lay[(t, "size")], t
        for n, _ in cls._fields_:
            assert getattr(cls, n).offset == lay[(t, n)], (t, n)
    assert ctypes.sizeof(S.Fields) == ctypes.sizeof(O.Fields)
    for (a, ta), (b, tb) in zip(S.Fields._fields_, O.Fields._fields_):
        assert a == b and ta == tb


def test_defaults_are_the_reference_defaults():
    f = S.default_fields()
    # bin/trapparameters.cpp:2-10, bin/ode.cpp:646-653
    assert (f.trap, f.kappa, f.Vdc, f.Vrf, f.r0, f.z0, f.freq_rf) == (1, 0.001, 5, 50, 0.04, 0.04, 50e5)
    assert (f.coulomb, f.coulomb_scale, f.tof) == (0, 1.0, 0)
    cfg = S.Config()
    S.lib().sbk_default_config(ctypes.byref(cfg))
    assert (cfg.rtol, cfg.atol, cfg.h0, cfg.adaptive) == (1e-7, 1e-7, 1e-9, 1)


def test_engine_fails_loudly_without_gpu():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    with pytest.raises(S.SimbucaError):
        S.Engine()


def test_synth_masses_match_reference_rule():
    amu = np.array([133.0, 85.0, 39.0, 1.00727638])
    assert np.array_equal(synth.mass_kg(amu), O.mass_kg(amu))


def test_synth_penning_frequencies():
    nu_c, nu_p, nu_m, nu_z = synth.penning_frequencies(133.0)
    assert nu_c == pytest.approx(808216.5, rel=1e-6)
    assert nu_m == pytest.approx(940.62, rel=1e-4)
    assert nu_p + nu_m == pytest.approx(nu_c, rel=1e-12)
    assert 2 * nu_p * nu_m == pytest.approx(nu_z ** 2, rel=1e-9)


def test_synth_workloads_shapes():
    w = synth.workload(3, n_override=2048)
    c = w.clouds[0]
    assert c.pos.shape == (2048, 3) and c.charge.dtype == np.int32
    assert c.pos.std() == pytest.approx(1.5e-3, rel=0.05)
    sig_v = np.sqrt(1.3806504e-23 * 300.0 / c.mass[0])
    assert c.vel.std() == pytest.approx(sig_v, rel=0.05)
    w4 = synth.workload(4, n_override=300)
    sizes = [cl.n for cl in w4.clouds]
    assert len(sizes) == 256 and min(sizes) >= 64 and max(sizes) <= 300
    w2 = synth.workload(2, n_override=4000)
    frac = np.mean(w2.clouds[0].mass_amu == 85.0)
    assert 0.45 < frac < 0.55


def test_cxx_caller_links_against_the_header_only():
    # built with g++ from include/simbuca_hip.h; loads the library, prints usage, no GPU work
    import subprocess
    exe = os.path.join(ROOT, "tests", "cxx", "abi_caller")
    assert os.path.exists(exe), "build tests/cxx (make -C tests/cxx)"
    p = subprocess.run([exe], capture_output=True, text=True, timeout=60)
    assert p.returncode == 2 and "usage" in p.stderr


@pytest.mark.parametrize("call,msg", [
    ("L.cunbody1_force(None, None, None, ctypes.c_double(0.0), None, -1, 0)",
     "cunbody1_force_begin: bad arguments"),
    ("L.cunbody1_force_end(None, None, None, ctypes.c_double(0.0), None, 4, 4)",
     "no matching cunbody1_force_begin"),
    ("N.initialize(0, False)", "n_max must be > 0"),
    ("N.simulation_Async(None, None, 8, 0, 0, 0)", "initialize() was not called"),
])
def test_reference_abi_errors_end_the_process(call, msg):
    # the reference backends' void entry points report errors by printing and
    # exit(EXIT_FAILURE) (bin/nbody_kernel.cu:37-47); these argument errors are
    # caught before any device call
    import subprocess
    import sys
    code = (f"import ctypes; L = ctypes.CDLL({S.LIB_PATH!r}); N = ctypes.CDLL({S.NBODY_LIB_PATH!r}); "
            f"{call}; print('returned')")
    p = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=120)
    assert p.returncode == 1, (p.returncode, p.stdout, p.stderr)
    assert msg in p.stderr and "returned" not in p.stdout
