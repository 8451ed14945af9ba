"""CPU oracle for the Simbuca force + Dormand-Prince path (test infrastructure).

Only tests/, ``__graft_entry__.smoke()`` and the CPU-baseline leg of
``bench.py`` import this package, and only as the checker.  The product
library (``simbuca_amd``) never calls it.

Two checkers live here:

* ``libsimbuca_oracle.so`` -- a plain-C FP64 restatement of the reference's
  force() (bin/force.cpp:34-175) and Dormand_Prince_5/error_succes/step
  (bin/ode.cpp:61-218,426-494), see ``simbuca_oracle.c``;
* ``_ref/ref_harness`` -- the reference's own CPU sources compiled in place
  (bin/*.cpp with -D__CPUNBODY_ON__) behind a small driver, when it was built;
* ``_ref/simbuca`` and ``_ref/ref_init`` -- the reference program itself and a
  driver of its InitCloud and mass tables (``make -C oracle refsim``).  They
  read the reference's mass tables from ``../libraries`` at start-up, so they
  run in a scratch directory whose parent links ``libraries`` to the
  reference's; nothing is written outside that scratch directory.
"""
from __future__ import annotations

import ctypes
import json
import os
import struct
import subprocess
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libsimbuca_oracle.so")
REF_HARNESS = os.path.join(HERE, "_ref", "ref_harness")
REF_SIMBUCA = os.path.join(HERE, "_ref", "simbuca")
REF_INIT = os.path.join(HERE, "_ref", "ref_init")
REF_LIBRARIES = os.path.join(os.environ.get("SIMBUCA_REFERENCE", "/root/reference"), "libraries")

TIME_FROZEN = 0
TIME_STAGE = 1


class Fields(ctypes.Structure):
    """Mirror of ``orc_fields`` (simbuca_oracle.h)."""

    _fields_ = [
        ("trap", ctypes.c_int),
        ("kappa", ctypes.c_double), ("Vdc", ctypes.c_double), ("Vrf", ctypes.c_double),
        ("r0", ctypes.c_double), ("z0", ctypes.c_double), ("freq_rf", ctypes.c_double),
        ("tof", ctypes.c_int), ("E_kick", ctypes.c_double),
        ("coulomb", ctypes.c_int), ("coulomb_scale", ctypes.c_double),
        ("V0", ctypes.c_double), ("d0", ctypes.c_double),
        ("Bz", ctypes.c_double),
        ("Ed", ctypes.c_double), ("fd", ctypes.c_double), ("phid", ctypes.c_double),
        ("Uq", ctypes.c_double), ("rq", ctypes.c_double), ("fq", ctypes.c_double),
        ("phiq", ctypes.c_double),
        ("delta", ctypes.c_double),
        ("ramp", ctypes.c_int), ("ramp_t0", ctypes.c_double), ("ramp_t1", ctypes.c_double),
        ("newVrf", ctypes.c_double), ("newVdc", ctypes.c_double),
        ("Uo", ctypes.c_double), ("ro", ctypes.c_double), ("fo", ctypes.c_double),
        ("phio", ctypes.c_double),
        ("rw_order", ctypes.c_int), ("Uw", ctypes.c_double), ("rw", ctypes.c_double),
        ("fw", ctypes.c_double), ("phiw", ctypes.c_double),
        ("Eax", ctypes.c_double), ("fax", ctypes.c_double), ("phiax", ctypes.c_double),
        ("fd_rate", ctypes.c_double), ("fq_rate", ctypes.c_double),
        ("B2", ctypes.c_double),
        ("K0", ctypes.c_double), ("p_mbar", ctypes.c_double), ("T_gas", ctypes.c_double),
        ("swift", ctypes.POINTER(ctypes.c_double)), ("swift_n", ctypes.c_int),
        ("swift_kind", ctypes.c_int), ("swift_t0", ctypes.c_double), ("swift_dt", ctypes.c_double),
        ("swift_r", ctypes.c_double),
        ("C4", ctypes.c_double), ("C6", ctypes.c_double),
        ("emap", ctypes.POINTER(ctypes.c_double)), ("emap_nr", ctypes.c_int),
        ("emap_nz", ctypes.c_int), ("emap_dr", ctypes.c_double), ("emap_z0", ctypes.c_double),
        ("emap_dz", ctypes.c_double), ("emap_scale", ctypes.c_double),
    ]


def fields(**kw) -> Fields:
    """Reference defaults (bin/trapparameters.cpp:2-10, bin/ode.cpp:646-653)
    with the extension terms off; keyword overrides."""
    f = Fields()
    f.trap = 1
    f.kappa, f.Vdc, f.Vrf, f.r0, f.z0, f.freq_rf = 0.001, 5.0, 50.0, 0.04, 0.04, 50e5
    f.tof, f.E_kick = 0, 0.0
    f.coulomb, f.coulomb_scale = 0, 1.0
    f.rq = 1.0
    f.ro, f.rw, f.rw_order, f.T_gas = 1.0, 1.0, 2, 273.15
    for k, v in kw.items():
        if not hasattr(f, k):
            raise KeyError(k)
        setattr(f, k, v)
    return f

def set_swift(f, samples, t0, dt, r, kind=1):
    """Attaches a SWIFT waveform to fields f (samples in V at t0 + k dt;
    kind 1: dipole along x, E = U/r; kind 2: quadrupole (U/r^2)(x, -y, 0)).
    The fields object keeps the sample array alive."""
    a = np.ascontiguousarray(samples, dtype=np.float64).ravel()
    f._swift_keep = a
    f.swift = a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    f.swift_n = int(a.size)
    f.swift_kind = int(kind)
    f.swift_t0, f.swift_dt, f.swift_r = float(t0), float(dt), float(r)
    return f


def set_emap(f, er, ez, dr, z0, dz, scale=1.0):
    """Attaches an axisymmetric field map to fields f: er, ez [nr, nz] in V/m
    at r = i dr, z = z0 + j dz; bilinear, 0 outside, times scale.  The
    fields object keeps the grid alive."""
    er = np.asarray(er, dtype=np.float64)
    ez = np.asarray(ez, dtype=np.float64)
    if er.ndim != 2 or er.shape != ez.shape:
        raise ValueError("er and ez must be [nr, nz] grids of one shape")
    g = np.ascontiguousarray(np.stack([er, ez], axis=-1)).ravel()
    f._emap_keep = g
    f.emap = g.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    f.emap_nr, f.emap_nz = int(er.shape[0]), int(er.shape[1])
    f.emap_dr, f.emap_z0, f.emap_dz, f.emap_scale = float(dr), float(z0), float(dz), float(scale)
    return f



_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle not built: {LIB_PATH} (run make -C oracle)")
        L = ctypes.CDLL(LIB_PATH)
        dp = ctypes.POINTER(ctypes.c_double)
        ip = ctypes.POINTER(ctypes.c_int)
        L.orc_mass_kg.restype = ctypes.c_double
        L.orc_mass_kg.argtypes = [ctypes.c_double]
        L.orc_force.restype = None
        L.orc_force.argtypes = [ctypes.c_int, dp, dp, dp, ip, ctypes.POINTER(Fields),
                                ctypes.c_double, dp]
        L.orc_coulomb_rows.restype = None
        L.orc_coulomb_rows.argtypes = [ctypes.c_int, dp, dp, ip, ctypes.c_double, ip,
                                       ctypes.c_int, dp]
        L.orc_cunbody_force.restype = None
        L.orc_cunbody_force.argtypes = [dp, dp, dp, ctypes.c_double, dp, ctypes.c_int,
                                        ctypes.c_int]
        L.orc_controller.restype = ctypes.c_int
        L.orc_controller.argtypes = [ctypes.c_double, dp, dp, dp, ip]
        L.orc_controller_beta.restype = ctypes.c_int
        L.orc_controller_beta.argtypes = [ctypes.c_double, ctypes.c_double, dp, dp, dp, ip]
        L.orc_set_beta.restype = None
        L.orc_set_beta.argtypes = [ctypes.c_void_p, ctypes.c_double]
        L.orc_create.restype = ctypes.c_void_p
        L.orc_create.argtypes = [ctypes.c_int, dp, dp, dp, ip, ctypes.POINTER(Fields),
                                 ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                 ctypes.c_double, ctypes.c_int, ctypes.c_int]
        L.orc_destroy.restype = None
        L.orc_destroy.argtypes = [ctypes.c_void_p]
        L.orc_set_fields.restype = None
        L.orc_set_fields.argtypes = [ctypes.c_void_p, ctypes.POINTER(Fields)]
        L.orc_dp5_attempt.restype = ctypes.c_double
        L.orc_dp5_attempt.argtypes = [ctypes.c_void_p]
        L.orc_step.restype = ctypes.c_int
        L.orc_step.argtypes = [ctypes.c_void_p]
        L.orc_advance.restype = ctypes.c_long
        L.orc_advance.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_long, dp,
                                  ctypes.c_long]
        L.orc_get_state.restype = None
        L.orc_get_state.argtypes = [ctypes.c_void_p, dp, dp, dp]
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _ip(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))


def _f64(a, shape=None):
    a = np.ascontiguousarray(a, dtype=np.float64)
    if shape is not None:
        a = a.reshape(shape)
    return a


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def mass_kg(mass_amu):
    """Ion mass in kg as bin/ion.cpp:3-6 computes it (long double product)."""
    m = np.asarray(mass_amu, dtype=np.float64)
    f = lib().orc_mass_kg
    return np.array([f(float(x)) for x in m.ravel()]).reshape(m.shape)


def force(pos, vel, mass, charge, f: Fields, t: float):
    """Restated force(): acceleration derivs[n][3]."""
    pos = _f64(pos, (-1, 3)); vel = _f64(vel, (-1, 3))
    mass = _f64(mass); charge = _i32(charge)
    n = pos.shape[0]
    out = np.zeros((n, 3))
    lib().orc_force(n, _dp(pos), _dp(vel), _dp(mass), _ip(charge), ctypes.byref(f), t, _dp(out))
    return out


def coulomb_rows(pos, mass, charge, coulomb_scale, rows):
    """f_j of force.cpp:150-170 for target rows (derivs_j -= f_j)."""
    pos = _f64(pos, (-1, 3)); mass = _f64(mass); charge = _i32(charge)
    rows = _i32(rows)
    out = np.zeros((rows.size, 3))
    lib().orc_coulomb_rows(pos.shape[0], _dp(pos), _dp(mass), _ip(charge),
                           float(coulomb_scale), _ip(rows), rows.size, _dp(out))
    return out


def cunbody_force(xj, mj, xi, eps2):
    xj = _f64(xj, (-1, 3)); xi = _f64(xi, (-1, 3)); mj = _f64(mj)
    out = np.zeros((xi.shape[0], 3))
    lib().orc_cunbody_force(_dp(xj), _dp(mj), _dp(xi), float(eps2), _dp(out),
                            xi.shape[0], xj.shape[0])
    return out


def controller(err, h, hnext, errold, reject, beta=0.0):
    """error_succes on a scalar error: returns (accept, h, hnext, errold, reject).
    beta != 0 is the reference's documented PI switch (bin/ode.cpp:427-432)."""
    H = ctypes.c_double(h); Hn = ctypes.c_double(hnext); Eo = ctypes.c_double(errold)
    R = ctypes.c_int(int(reject))
    ok = lib().orc_controller_beta(float(err), float(beta), ctypes.byref(H), ctypes.byref(Hn),
                                   ctypes.byref(Eo), ctypes.byref(R))
    return bool(ok), H.value, Hn.value, Eo.value, bool(R.value)


class Cloud:
    """Restated IonCloud + _ode_vars pair driven by step()/move_particles()."""

    def __init__(self, pos, vel, mass, charge, f: Fields, lifetime=0.0, h0=1e-9,
                 rtol=1e-7, atol=1e-7, adaptive=True, time_mode=TIME_FROZEN, beta=0.0):
        pos = _f64(pos, (-1, 3)); vel = _f64(vel, (-1, 3))
        mass = _f64(mass); charge = _i32(charge)
        self.n = pos.shape[0]
        self._h = lib().orc_create(self.n, _dp(pos), _dp(vel), _dp(mass), _ip(charge),
                                   ctypes.byref(f), lifetime, h0, rtol, atol,
                                   1 if adaptive else 0, time_mode)
        self._fields = f       # the C copy points at its SWIFT samples
        if beta:
            lib().orc_set_beta(self._h, float(beta))

    def __del__(self):
        if getattr(self, "_h", None):
            lib().orc_destroy(self._h)
            self._h = None

    def set_fields(self, f: Fields):
        lib().orc_set_fields(self._h, ctypes.byref(f))
        self._fields = f

    def attempt(self):
        return lib().orc_dp5_attempt(self._h)

    def step(self):
        if lib().orc_step(self._h) != 0:
            raise RuntimeError("stepsize underflow created")

    def advance(self, t_end, max_steps=1 << 40, hist_cap=0):
        hist = np.zeros(max(hist_cap, 1))
        k = lib().orc_advance(self._h, float(t_end), int(max_steps), _dp(hist), int(hist_cap))
        if k < 0:
            raise RuntimeError("stepsize underflow created")
        return int(k), hist[:min(k, hist_cap)]

    def state(self):
        pos = np.zeros((self.n, 3)); vel = np.zeros((self.n, 3)); s = np.zeros(8)
        lib().orc_get_state(self._h, _dp(pos), _dp(vel), _dp(s))
        keys = ["lifetime", "h", "hnext", "errold", "reject", "steps", "attempts", "last_err"]
        return pos, vel, dict(zip(keys, s.tolist()))


# ---------------------------------------------------------------------------
# the reference's own CPU path (oracle/_ref), when it has been built
# ---------------------------------------------------------------------------

def ref_available() -> bool:
    return os.path.exists(REF_HARNESS) and os.access(REF_HARNESS, os.X_OK)


def _ref_input(mode, mass_amu, charge, pos, vel, f: Fields, lifetime, h0=1e-9, rtol=1e-7,
               atol=1e-7, adaptive=True, nsteps=0):
    n = len(mass_amu)
    hdr = struct.pack("<8i", 0x53424B31, mode, n, nsteps, int(f.trap), int(f.tof),
                      int(f.coulomb), 1 if adaptive else 0)
    par = struct.pack("<12d", lifetime, h0, rtol, atol, f.kappa, f.Vdc, f.Vrf, f.r0, f.z0,
                      f.freq_rf, f.E_kick, f.coulomb_scale)
    return (hdr + par + _f64(mass_amu).tobytes() + _i32(charge).tobytes()
            + _f64(pos, (-1, 3)).tobytes() + _f64(vel, (-1, 3)).tobytes())


def _run_ref(payload: bytes, want_output: bool, timeout=600):
    with tempfile.TemporaryDirectory() as d:
        fin = os.path.join(d, "in.bin"); fout = os.path.join(d, "out.bin")
        with open(fin, "wb") as fh:
            fh.write(payload)
        p = subprocess.run([REF_HARNESS, fin, fout if want_output else "-"],
                           capture_output=True, text=True, timeout=timeout)
        if p.returncode != 0:
            raise RuntimeError(f"ref_harness failed: {p.stderr}")
        if want_output:
            with open(fout, "rb") as fh:
                return fh.read()
        return p.stdout


def ref_force(mass_amu, charge, pos, vel, f: Fields, lifetime):
    """The reference's force() on the given state (only trap/TOF/Coulomb terms)."""
    raw = _run_ref(_ref_input(1, mass_amu, charge, pos, vel, f, lifetime), True)
    return np.frombuffer(raw, dtype=np.float64).reshape(-1, 3).copy()


def ref_steps(mass_amu, charge, pos, vel, f: Fields, lifetime, h0, rtol, atol, adaptive,
              nsteps):
    """nsteps calls of the reference step(); returns pos, vel, scalars, hist."""
    n = len(mass_amu)
    raw = _run_ref(_ref_input(2, mass_amu, charge, pos, vel, f, lifetime, h0, rtol, atol,
                              adaptive, nsteps), True)
    a = np.frombuffer(raw, dtype=np.float64)
    pos_o = a[:3 * n].reshape(n, 3).copy(); vel_o = a[3 * n:6 * n].reshape(n, 3).copy()
    sc = a[6 * n:6 * n + 8].copy(); hist = a[6 * n + 8:].reshape(-1, 2).copy()
    keys = ["lifetime", "h", "hnext", "errold", "reject", "countstep", "err", "_"]
    return pos_o, vel_o, dict(zip(keys, sc.tolist())), hist


def ref_output(mass_amu, charge, pos, vel, lifetime, prints, out_dir, per_particle=False):
    """The reference's own output files (IonCloud::PrintParticles/PrintCloud/
    Delete): a print at `lifetime`, then `prints` times pos += 1e-7 vel,
    vel *= 0.75, lifetime += 1e-7 and a print, then Delete().  Writes under
    out_dir with prefix "ref"; returns {file name: text}.  Ion k is named
    "%g" % mass_amu[k] + "u"."""
    prefix = os.path.join(out_dir, "ref")
    payload = _ref_input(6 if per_particle else 5, mass_amu, charge, pos, vel, fields(), lifetime,
                         nsteps=prints)
    fin = os.path.join(out_dir, "in.bin")
    with open(fin, "wb") as fh:
        fh.write(payload)
    p = subprocess.run([REF_HARNESS, fin, prefix], capture_output=True, text=True, timeout=120,
                       cwd=out_dir)
    if p.returncode != 0:
        raise RuntimeError(f"ref_harness failed: {p.stderr}")
    out = {}
    for name in sorted(os.listdir(out_dir)):
        if name.startswith("ref_") and name.endswith(".txt"):
            with open(os.path.join(out_dir, name)) as fh:
                out[name[len("ref"):]] = fh.read()
    return out


def ref_time(mode, mass_amu, charge, pos, vel, f: Fields, lifetime, calls, h0=1e-9,
             rtol=1e-7, atol=1e-7, timeout=600):
    """Time `calls` reference force() (mode 3) or step() (mode 4) calls."""
    out = _run_ref(_ref_input(mode, mass_amu, charge, pos, vel, f, lifetime, h0, rtol, atol,
                              True, calls), False, timeout=timeout)
    return json.loads(out.strip().splitlines()[-1])


# ---------------------------------------------------------------------------
# the reference program and its initializer (oracle/_ref/simbuca, ref_init)
# ---------------------------------------------------------------------------

def refsim_available() -> bool:
    return (all(os.path.exists(p) and os.access(p, os.X_OK) for p in (REF_SIMBUCA, REF_INIT))
            and os.path.isdir(REF_LIBRARIES))


def _scratch(d):
    """d/libraries -> the reference's tables; d/run is the working directory."""
    os.symlink(REF_LIBRARIES, os.path.join(d, "libraries"))
    run = os.path.join(d, "run")
    os.mkdir(run)
    return run


def ref_init_cloud(n, seed, energy_eV, semiaxis, offset, mass_amu):
    """The reference's InitCloud: (pos [n,3], vel [n,3]) exactly as it draws them."""
    e = [float(x) for x in np.atleast_1d(energy_eV)]
    args = [REF_INIT, "init", str(n), str(seed), str(len(e))] + ["%.17g" % x for x in e]
    args += ["%.17g" % float(x) for x in semiaxis] + ["%.17g" % float(x) for x in offset]
    args += ["%.17g" % float(x) for x in mass_amu]
    with tempfile.TemporaryDirectory() as d:
        p = subprocess.run(args, capture_output=True, text=True, timeout=120, cwd=_scratch(d))
    if p.returncode != 0:
        raise RuntimeError(f"ref_init failed: {p.stderr}")
    rows = [ln.split() for ln in p.stdout.splitlines()]
    a = np.array([[float(x) for x in r] for r in rows if len(r) == 6])
    if a.shape != (n, 6):
        raise RuntimeError("ref_init printed %s rows" % (a.shape,))
    return a[:, :3].copy(), a[:, 3:].copy()


def ref_masses(names):
    """{name: (table, mass in u)} from the reference's mass tables
    (table "ame", "pdg" or "none" with mass nan)."""
    with tempfile.TemporaryDirectory() as d:
        p = subprocess.run([REF_INIT, "mass"] + list(names), capture_output=True, text=True,
                           timeout=120, cwd=_scratch(d))
    if p.returncode != 0:
        raise RuntimeError(f"ref_init failed: {p.stderr}")
    out = {}
    for ln in p.stdout.splitlines():
        r = ln.split()
        if len(r) >= 2 and r[1] in ("ame", "pdg", "none"):
            out[r[0]] = (r[1], float(r[2]) if r[1] != "none" else float("nan"))
    return out


def ref_run_ini(ini_text, timeout=600):
    """Runs the reference program on a simulation file; returns {path relative
    to the working directory: text} of the .txt files it wrote."""
    with tempfile.TemporaryDirectory() as d:
        run = _scratch(d)
        with open(os.path.join(run, "sim.ini"), "w") as fh:
            fh.write(ini_text)
        p = subprocess.run([REF_SIMBUCA, "sim.ini"], capture_output=True, text=True,
                           timeout=timeout, cwd=run)
        if p.returncode != 0:
            raise RuntimeError(f"reference program failed: {p.stdout[-2000:]} {p.stderr}")
        out = {}
        for root, _, files in os.walk(run):
            for name in files:
                if name.endswith(".txt"):
                    path = os.path.join(root, name)
                    with open(path) as fh:
                        out[os.path.relpath(path, run)] = fh.read()
        return out


def ref_ini(path, queries):
    """The reference's INIReader on a file: (sections, [(fields of section,
    Get(section, name) or None)] for each (section, name) query)."""
    args = [REF_INIT, "ini", os.path.abspath(path)]
    for sec, name in queries:
        args += [sec, name]
    with tempfile.TemporaryDirectory() as d:
        p = subprocess.run(args, capture_output=True, timeout=120, cwd=_scratch(d))
    if p.returncode != 0:
        raise RuntimeError(f"ref_init failed: {p.stderr!r}")

    def unhex(h):
        return "" if h == "=" else bytes.fromhex(h).decode("latin-1")
    sections, out, fields = [], [], []
    for ln in p.stdout.decode().splitlines():
        kind, _, val = ln.partition(" ")
        if kind == "S":
            sections.append(unhex(val))
        elif kind == "F":
            fields.append(unhex(val))
        elif kind == "V":
            out.append((fields, None if val == "-" else unhex(val)))
            fields = []
    return sections, out
