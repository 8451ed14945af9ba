"""simbuca_amd -- MI355X force and Dormand-Prince 5(4) engine for ion clouds.

Python binding of ``libsimbuca_hip.so`` (C ABI in ``include/simbuca_hip.h``).
The compute path is the HIP library; this module only marshals arrays.  If
the library has not been built it raises instead of falling back to any CPU
implementation.

Names follow the reference (mgrau/Simbuca): an ``Engine`` holds one or more
ion clouds (``IonCloud``, bin/ioncloud.h:17-71) and advances them with the
device-resident equivalent of ``step()``/``move_particles()``
(bin/ode.cpp:459-494, bin/ion_fly.cpp:17-122); ``Fields`` mirrors the trap
parameters of bin/trapparameters.h:6-27 plus the Penning-trap terms.
"""
from __future__ import annotations

import ctypes
import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libsimbuca_hip.so")
NBODY_LIB_PATH = os.path.join(HERE, "libsimbuca_nbody.so")

FP64_EXACT, FP64, MIXED, MIXED_KAHAN, FP64_RSQ = 0, 1, 2, 3, 4
PRECISIONS = {"fp64_exact": FP64_EXACT, "fp64": FP64, "mixed": MIXED, "mixed_kahan": MIXED_KAHAN,
              "fp64_rsq": FP64_RSQ}
TIME_FROZEN, TIME_STAGE = 0, 1

SBK_EUNDERFLOW = -4

# bin/globals.h:6-15
EL_CHARGE = 1.60217733e-19
AMU = 1.6605402e-27
KE = 8.987551787e9
KB = 1.3806504e-23


class SimbucaError(RuntimeError):
    pass


class StepsizeUnderflow(SimbucaError):
    """The reference throws "stepsize underflow created" (bin/ode.cpp:482-483)."""


class Fields(ctypes.Structure):
    """``sbk_fields`` of include/simbuca_hip.h."""

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

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class Config(ctypes.Structure):
    _fields_ = [
        ("precision", ctypes.c_int), ("time_mode", ctypes.c_int), ("adaptive", ctypes.c_int),
        ("use_graph", ctypes.c_int), ("timing", ctypes.c_int),
        ("rtol", ctypes.c_double), ("atol", ctypes.c_double), ("h0", ctypes.c_double),
        ("spatial_order", ctypes.c_int), ("reorder_interval", ctypes.c_int),
        ("beta", ctypes.c_double), ("persistent", ctypes.c_int), ("persistent_spin", ctypes.c_int),
    ]


class CloudStatus(ctypes.Structure):
    _fields_ = [
        ("lifetime", ctypes.c_double), ("h", ctypes.c_double), ("hnext", ctypes.c_double),
        ("errold", ctypes.c_double), ("last_err", ctypes.c_double), ("t_end", ctypes.c_double),
        ("reject", ctypes.c_int), ("active", ctypes.c_int), ("accepted_last", ctypes.c_int),
        ("error", ctypes.c_int),
        ("steps", ctypes.c_longlong), ("attempts", ctypes.c_longlong),
        ("n", ctypes.c_int), ("slices", ctypes.c_int),
        ("coulomb_evals", ctypes.c_longlong), ("t_prev", ctypes.c_double),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class Stats(ctypes.Structure):
    _fields_ = [
        ("coulomb_ms", ctypes.c_double), ("coulomb_launches", ctypes.c_longlong),
        ("pair_interactions", ctypes.c_double), ("attempts", ctypes.c_longlong),
        ("steps", ctypes.c_longlong), ("ion_steps", ctypes.c_double),
        ("reorders", ctypes.c_longlong), ("persistent_launches", ctypes.c_longlong),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


_lib = None

# every symbol include/simbuca_hip.h declares
EXPORTS = [
    "cunbody1_force", "cunbody1_force_begin", "cunbody1_force_end", "sbk_cunbody_set_precision",
    "sbk_cunbody_force_static",
    "sbk_default_config", "sbk_default_fields", "sbk_create", "sbk_destroy", "sbk_add_cloud",
    "sbk_set_fields", "sbk_set_step", "sbk_advance", "sbk_advance_until", "sbk_step", "sbk_attempts",
    "sbk_synchronize", "sbk_reorder", "sbk_set_crossing_plane", "sbk_take_crossings", "sbk_get_state", "sbk_get_status", "sbk_get_stats", "sbk_reset_stats",
    "sbk_eval_force", "sbk_eval_coulomb", "sbk_set_launch", "sbk_set_coulomb_kernel", "sbk_get_launch",
    "sbk_time_coulomb", "sbk_autotune",
    "sbk_load_tuning", "sbk_spatial_order", "sbk_writer_open", "sbk_writer_particles",
    "sbk_writer_cloud", "sbk_writer_set_ion", "sbk_writer_x", "sbk_writer_close", "sbk_import", "sbk_init_cloud", "sbk_mass_table_open",
    "sbk_mass_table_lookup", "sbk_mass_table_close", "sbk_parse_species", "sbk_run_ini", "sbk_ini_describe", "sbk_ini_value", "sbk_ini_names",
    "sbk_last_error", "sbk_device_info",
]

# every symbol include/simbuca_nbody.h declares (libsimbuca_nbody.so)
NBODY_EXPORTS = ["initialize", "finalize", "simulation_Async", "simulation_Async_begin",
                 "simulation_Async_end"]


def lib():
    """Loads libsimbuca_hip.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise SimbucaError(
            f"{LIB_PATH} is missing: build it with `make -C simbuca_amd/csrc` "
            "(or __graft_entry__.build()); there is no CPU fallback")
    L = ctypes.CDLL(LIB_PATH)
    dp = ctypes.POINTER(ctypes.c_double)
    ip = ctypes.POINTER(ctypes.c_int)
    vp = ctypes.c_void_p
    c_int, c_dbl, c_ll = ctypes.c_int, ctypes.c_double, ctypes.c_longlong
    sig = {
        "cunbody1_force": (None, [dp, dp, dp, c_dbl, dp, c_int, c_int]),
        "cunbody1_force_begin": (None, [dp, dp, dp, c_dbl, dp, c_int, c_int]),
        "cunbody1_force_end": (None, [dp, dp, dp, c_dbl, dp, c_int, c_int]),
        "sbk_cunbody_set_precision": (c_int, [c_int]),
        "sbk_cunbody_force_static": (c_int, [dp, dp, dp, c_dbl, dp, c_int, c_int]),
        "sbk_default_config": (None, [ctypes.POINTER(Config)]),
        "sbk_default_fields": (None, [ctypes.POINTER(Fields)]),
        "sbk_create": (c_int, [ctypes.POINTER(Config), ctypes.POINTER(vp)]),
        "sbk_destroy": (c_int, [vp]),
        "sbk_add_cloud": (c_int, [vp, c_int, dp, dp, dp, ip, ctypes.POINTER(Fields), c_dbl, ip]),
        "sbk_set_fields": (c_int, [vp, c_int, ctypes.POINTER(Fields)]),
        "sbk_set_step": (c_int, [vp, c_int, c_dbl]),
        "sbk_advance": (c_int, [vp, c_dbl, c_ll]),
        "sbk_advance_until": (c_int, [vp, c_dbl, c_ll]),
        "sbk_step": (c_int, [vp]),
        "sbk_attempts": (c_int, [vp, c_int]),
        "sbk_synchronize": (c_int, [vp]),
        "sbk_reorder": (c_int, [vp]),
        "sbk_set_crossing_plane": (c_int, [vp, c_int, c_int, c_dbl, ip]),
        "sbk_take_crossings": (c_int, [vp, c_int, ip, ip, ip, ctypes.POINTER(c_ll), dp, dp, dp]),
        "sbk_get_state": (c_int, [vp, c_int, dp, dp]),
        "sbk_get_status": (c_int, [vp, c_int, ctypes.POINTER(CloudStatus)]),
        "sbk_get_stats": (c_int, [vp, ctypes.POINTER(Stats)]),
        "sbk_reset_stats": (c_int, [vp]),
        "sbk_eval_force": (c_int, [vp, c_int, dp]),
        "sbk_eval_coulomb": (c_int, [vp, c_int, c_int, dp]),
        "sbk_set_launch": (c_int, [vp, c_int, c_int, c_int, c_int]),
        "sbk_set_coulomb_kernel": (c_int, [vp, c_int]),
        "sbk_get_launch": (c_int, [vp, c_int, ip, ip, ip, ip, ip]),
        "sbk_time_coulomb": (c_int, [vp, c_int, c_int, dp]),
        "sbk_autotune": (c_int, [vp, c_int, ip, ip, ip, ip, ip, ctypes.c_char_p]),
        "sbk_load_tuning": (c_int, [vp, c_int, ctypes.c_char_p]),
        "sbk_spatial_order": (c_int, [c_int, dp, ip]),
        "sbk_writer_open": (c_int, [ctypes.c_char_p, c_int, ctypes.POINTER(ctypes.c_char_p), dp, c_int,
                                    c_int, ctypes.POINTER(vp)]),
        "sbk_writer_particles": (c_int, [vp, c_dbl, dp, dp]),
        "sbk_writer_cloud": (c_int, [vp, c_dbl, dp, dp]),
        "sbk_writer_set_ion": (c_int, [vp, c_int, ctypes.c_char_p, c_dbl]),
        "sbk_writer_x": (c_int, [vp, c_int, c_dbl, dp, dp]),
        "sbk_writer_close": (c_int, [vp, c_dbl, dp, dp]),
        "sbk_import": (c_int, [ctypes.c_char_p, c_int, ip, ctypes.c_char_p, c_int, dp, dp, dp]),
        "sbk_init_cloud": (c_int, [c_int, ctypes.c_ulong, dp, dp, c_int, dp, dp, dp, dp]),
        "sbk_mass_table_open": (c_int, [ctypes.c_char_p, ctypes.c_char_p, ctypes.POINTER(vp)]),
        "sbk_mass_table_lookup": (c_int, [vp, ctypes.c_char_p, dp]),
        "sbk_mass_table_close": (None, [vp]),
        "sbk_parse_species": (c_int, [ctypes.c_char_p, vp, dp, ip]),
        "sbk_run_ini": (c_int, [ctypes.c_char_p, vp, c_int, c_int, c_int]),
        "sbk_ini_describe": (c_int, [ctypes.c_char_p, vp, ctypes.c_char_p, c_int]),
        "sbk_ini_value": (c_int, [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p, c_int]),
        "sbk_ini_names": (c_int, [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p, c_int]),
        "sbk_last_error": (ctypes.c_char_p, []),
        "sbk_device_info": (c_int, [ctypes.c_char_p, c_int, ip, ip]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


_nbody = None


def nbody_lib():
    """Loads libsimbuca_nbody.so (the reference's NBODY interface); raises if
    it has not been built."""
    global _nbody
    if _nbody is not None:
        return _nbody
    lib()   # its Coulomb kernels live in libsimbuca_hip.so
    if not os.path.exists(NBODY_LIB_PATH):
        raise SimbucaError(f"{NBODY_LIB_PATH} is missing: build it with `make -C simbuca_amd/csrc`")
    L = ctypes.CDLL(NBODY_LIB_PATH)
    vp, c_int = ctypes.c_void_p, ctypes.c_int
    sig = {
        "initialize": (None, [c_int, ctypes.c_bool]),
        "finalize": (ctypes.c_float, []),
        "simulation_Async": (None, [vp, vp, c_int, c_int, c_int, c_int]),
        "simulation_Async_begin": (None, [vp, vp, c_int, c_int, c_int, c_int]),
        "simulation_Async_end": (None, [vp, vp, c_int, c_int, c_int, c_int]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _nbody = L
    return L


def _check(rc):
    if rc != 0:
        msg = lib().sbk_last_error().decode(errors="replace")
        if rc == SBK_EUNDERFLOW:
            raise StepsizeUnderflow(msg)
        raise SimbucaError(f"libsimbuca_hip error {rc}: {msg}")


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _ip(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))


def default_fields(**kw) -> Fields:
    """Reference defaults (bin/trapparameters.cpp:2-10) with keyword overrides."""
    f = Fields()
    lib().sbk_default_fields(ctypes.byref(f))
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



def spatial_order(pos):
    """The order a spatial_order engine keeps a cloud's ions in on the device
    (host computation, no GPU needed): slot i holds ion order[i]."""
    pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1, 3)
    order = np.empty(pos.shape[0], dtype=np.int32)
    _check(lib().sbk_spatial_order(pos.shape[0], _dp(pos), _ip(order)))
    return order


class Writer:
    """The reference's output files (IonCloud printing) for one cloud: see
    sbk_writer_open.  Host only."""

    def __init__(self, prefix, names, mass_amu, per_particle_files=False, digits=6):
        names = [str(x) for x in names]
        mass = np.ascontiguousarray(mass_amu, dtype=np.float64).ravel()
        if len(names) != mass.size:
            raise ValueError("one name and one mass per ion")
        self.n = mass.size
        arr = (ctypes.c_char_p * max(1, self.n))(*[x.encode() for x in names])
        h = ctypes.c_void_p()
        _check(lib().sbk_writer_open(str(prefix).encode(), self.n, arr, _dp(mass),
                                     1 if per_particle_files else 0, int(digits), ctypes.byref(h)))
        self._h = h

    def _state(self, pos, vel):
        pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1, 3)
        vel = np.ascontiguousarray(vel, dtype=np.float64).reshape(-1, 3)
        if pos.shape[0] != self.n or vel.shape[0] != self.n:
            raise ValueError("state does not match the writer's ions")
        return pos, vel

    def particles(self, lifetime, pos, vel):
        pos, vel = self._state(pos, vel)
        _check(lib().sbk_writer_particles(self._h, float(lifetime), _dp(pos), _dp(vel)))

    def cloud(self, lifetime, pos, vel):
        pos, vel = self._state(pos, vel)
        _check(lib().sbk_writer_cloud(self._h, float(lifetime), _dp(pos), _dp(vel)))

    def set_ion(self, k, name, mass_kg):
        """Renames ion k and sets the mass (kg) of its printed energy."""
        _check(lib().sbk_writer_set_ion(self._h, int(k), str(name).encode(), float(mass_kg)))

    def x(self, k, lifetime, pos, vel):
        """The reference's PrintX line for ion k (the first call writes the header only)."""
        pos, vel = self._state(pos, vel)
        _check(lib().sbk_writer_x(self._h, int(k), float(lifetime), _dp(pos), _dp(vel)))

    def close(self, lifetime=0.0, pos=None, vel=None):
        if not self._h:
            return
        h, self._h = self._h, None
        if pos is None:
            _check(lib().sbk_writer_close(h, float(lifetime), None, None))
        else:
            pos, vel = self._state(pos, vel)
            _check(lib().sbk_writer_close(h, float(lifetime), _dp(pos), _dp(vel)))

    def __del__(self):
        if getattr(self, "_h", None):
            lib().sbk_writer_close(self._h, 0.0, None, None)


def import_snapshot(prefix, capacity=None, name_len=64):
    """Last complete print of <prefix>_pAll.txt: (names, pos [m], vel [m/s], lifetime [s])."""
    n = ctypes.c_int()
    if capacity is None:
        _check(lib().sbk_import(str(prefix).encode(), 0, ctypes.byref(n), None, name_len, None,
                                None, None))
        capacity = n.value
    names = ctypes.create_string_buffer(capacity * name_len)
    pos = np.zeros((capacity, 3)); vel = np.zeros((capacity, 3)); t = ctypes.c_double()
    _check(lib().sbk_import(str(prefix).encode(), capacity, ctypes.byref(n), names, name_len,
                            _dp(pos), _dp(vel), ctypes.byref(t)))
    k = n.value
    raw = names.raw
    out = [raw[i * name_len:(i + 1) * name_len].split(b"\0", 1)[0].decode() for i in range(k)]
    return out, pos[:k].copy(), vel[:k].copy(), t.value


def init_cloud(n, mass_amu, energy_eV, semiaxis=(1e-3, 1e-3, 1e-3), offset=(0.0, 0.0, 0.0),
               seed=0):
    """The reference's InitCloud (host, no GPU): (pos [n,3] m, vel [n,3] m/s).
    energy_eV is one energy, or (radial, axial) for separate vz speeds."""
    mass = np.ascontiguousarray(np.broadcast_to(np.asarray(mass_amu, dtype=np.float64), (n,)))
    e = np.atleast_1d(np.asarray(energy_eV, dtype=np.float64)).copy()
    semi = np.ascontiguousarray(semiaxis, dtype=np.float64)
    off = np.ascontiguousarray(offset, dtype=np.float64)
    if semi.size != 3 or off.size != 3:
        raise ValueError("semiaxis and offset have three components")
    pos = np.empty((n, 3)); vel = np.empty((n, 3))
    _check(lib().sbk_init_cloud(int(n), int(seed), _dp(semi), _dp(off), int(e.size), _dp(e),
                                _dp(mass), _dp(pos), _dp(vel)))
    return pos, vel


class MassTable:
    """The reference's AME and PDG mass tables (either path may be None)."""

    def __init__(self, ame_path=None, pdg_path=None):
        h = ctypes.c_void_p()
        enc = lambda p: None if p is None else str(p).encode()
        _check(lib().sbk_mass_table_open(enc(ame_path), enc(pdg_path), ctypes.byref(h)))
        self._h = h

    def mass(self, name):
        m = ctypes.c_double()
        _check(lib().sbk_mass_table_lookup(self._h, str(name).encode(), ctypes.byref(m)))
        return m.value

    def close(self):
        if getattr(self, "_h", None):
            lib().sbk_mass_table_close(self._h)
            self._h = None

    __del__ = close


def parse_species(name, table=None):
    """(mass in u, charge) of a [composition] name."""
    m = ctypes.c_double(); q = ctypes.c_int()
    _check(lib().sbk_parse_species(str(name).encode(), table._h if table else None,
                                   ctypes.byref(m), ctypes.byref(q)))
    return m.value, q.value


def describe_ini(path, table=None):
    """The settings sbk_run_ini would run, as a dict (host only)."""
    buf = ctypes.create_string_buffer(1 << 20)
    _check(lib().sbk_ini_describe(str(path).encode(), table._h if table else None, buf, len(buf)))
    return json.loads(buf.value.decode())


def ini_value(path, section, name):
    """The raw value of [section] name as the reference's INIReader gives it,
    or None when it is not set."""
    buf = ctypes.create_string_buffer(1 << 16)
    rc = lib().sbk_ini_value(str(path).encode(), section.encode(), name.encode(), buf, len(buf))
    if rc == -1:      # SBK_EINVAL: not set
        return None
    _check(rc)
    return buf.value.decode()


def ini_names(path, section=None):
    """Sections (section None) or the field names of a section, sorted."""
    buf = ctypes.create_string_buffer(1 << 16)
    _check(lib().sbk_ini_names(str(path).encode(), None if section is None else section.encode(),
                               buf, len(buf)))
    return buf.value.decode().split("\n")[:-1]   # one per line, each ended by '\n' (names may be empty)


def run_ini(path, table=None, precision=FP64_EXACT, time_mode=TIME_FROZEN, verbose=False):
    """Runs a reference simulation file on the GPU (see sbk_run_ini); the
    output files go where its [output] prefix says, relative to the
    working directory."""
    _check(lib().sbk_run_ini(str(path).encode(), table._h if table else None, int(precision),
                             int(time_mode), 1 if verbose else 0))


def device_info():
    name = ctypes.create_string_buffer(128)
    cus = ctypes.c_int(); clk = ctypes.c_int()
    _check(lib().sbk_device_info(name, 128, ctypes.byref(cus), ctypes.byref(clk)))
    return {"name": name.value.decode(), "cus": cus.value, "clock_mhz": clk.value}


class Engine:
    """Device-resident clouds advanced by Dormand-Prince 5(4) on the GPU."""

    def __init__(self, precision="mixed", time_mode=TIME_FROZEN, adaptive=True, use_graph=True,
                 timing=False, rtol=1e-7, atol=1e-7, h0=1e-9, spatial_order=None,
                 reorder_interval=None, beta=0.0, persistent=0, persistent_spin=0):
        L = lib()
        cfg = Config()
        L.sbk_default_config(ctypes.byref(cfg))
        cfg.precision = PRECISIONS[precision] if isinstance(precision, str) else int(precision)
        cfg.time_mode = int(time_mode)
        cfg.adaptive = 1 if adaptive else 0
        cfg.use_graph = 1 if use_graph else 0
        cfg.timing = 1 if timing else 0
        cfg.rtol, cfg.atol, cfg.h0 = float(rtol), float(atol), float(h0)
        if spatial_order is not None:
            cfg.spatial_order = 1 if spatial_order else 0
        if reorder_interval is not None:
            cfg.reorder_interval = int(reorder_interval)
        cfg.beta = float(beta)
        cfg.persistent = int(persistent)
        cfg.persistent_spin = int(persistent_spin)
        h = ctypes.c_void_p()
        _check(L.sbk_create(ctypes.byref(cfg), ctypes.byref(h)))
        self._h = h
        self.config = cfg
        self.precision = next(k for k, v in PRECISIONS.items() if v == cfg.precision)
        self.sizes = []

    def close(self):
        if getattr(self, "_h", None):
            lib().sbk_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def add_cloud(self, pos, vel, mass, charge, fields: Fields, lifetime=0.0) -> int:
        pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1, 3)
        vel = np.ascontiguousarray(vel, dtype=np.float64).reshape(-1, 3)
        mass = np.ascontiguousarray(mass, dtype=np.float64).ravel()
        charge = np.ascontiguousarray(charge, dtype=np.int32).ravel()
        n = pos.shape[0]
        if vel.shape[0] != n or mass.size != n or charge.size != n:
            raise ValueError("pos, vel, mass and charge must describe the same ions")
        cid = ctypes.c_int()
        _check(lib().sbk_add_cloud(self._h, n, _dp(pos), _dp(vel), _dp(mass), _ip(charge),
                                   ctypes.byref(fields), float(lifetime), ctypes.byref(cid)))
        self.sizes.append(n)
        return cid.value

    def set_fields(self, cloud, fields: Fields):
        _check(lib().sbk_set_fields(self._h, cloud, ctypes.byref(fields)))

    def set_step(self, cloud, h):
        _check(lib().sbk_set_step(self._h, cloud, float(h)))

    def set_launch(self, targets_per_thread=0, source_groups=4, slices=0, local_coords=-1):
        _check(lib().sbk_set_launch(self._h, int(targets_per_thread), int(source_groups),
                                    int(slices), int(local_coords)))

    def set_coulomb_kernel(self, kernel):
        """'auto', 'tiled' or 'stream' (see sbk_set_coulomb_kernel)."""
        k = {"auto": -1, "tiled": 0, "stream": 1}[kernel] if isinstance(kernel, str) else int(kernel)
        _check(lib().sbk_set_coulomb_kernel(self._h, k))

    def advance(self, duration, max_attempts=1 << 62):
        _check(lib().sbk_advance(self._h, float(duration), int(max_attempts)))

    def advance_until(self, t_end, max_attempts=1 << 62):
        _check(lib().sbk_advance_until(self._h, float(t_end), int(max_attempts)))

    def step(self):
        _check(lib().sbk_step(self._h))

    def attempts(self, count, open_ended=True):
        """Enqueues `count` attempts without waiting.  With open_ended, first
        makes every cloud active with no end time (sbk_advance(inf, 0))."""
        if open_ended:
            _check(lib().sbk_advance(self._h, float("inf"), 0))
        _check(lib().sbk_attempts(self._h, int(count)))

    def synchronize(self):
        _check(lib().sbk_synchronize(self._h))

    def reorder(self):
        _check(lib().sbk_reorder(self._h))

    def set_crossing_plane(self, cloud, x_plane, enable=True):
        """Records the ions of `cloud` that cross x = x_plane in each accepted
        step (the reference's print_at_x); returns the device list's capacity."""
        cap = ctypes.c_int()
        _check(lib().sbk_set_crossing_plane(self._h, cloud, 1 if enable else 0, float(x_plane),
                                            ctypes.byref(cap)))
        return cap.value

    def take_crossings(self):
        """The recorded crossings ordered by cloud, step and ion (the list is
        emptied): dict of arrays cloud, ion, step, t, pos (n, 3), vel (n, 3)."""
        L = lib()
        n = ctypes.c_int()
        _check(L.sbk_take_crossings(self._h, 0, ctypes.byref(n), None, None, None, None, None, None))
        k = n.value
        out = {"cloud": np.zeros(k, np.int32), "ion": np.zeros(k, np.int32),
               "step": np.zeros(k, np.int64), "t": np.zeros(k),
               "pos": np.zeros((k, 3)), "vel": np.zeros((k, 3))}
        _check(L.sbk_take_crossings(
            self._h, max(k, 1), ctypes.byref(n), _ip(out["cloud"]), _ip(out["ion"]),
            out["step"].ctypes.data_as(ctypes.POINTER(ctypes.c_longlong)), _dp(out["t"]),
            _dp(out["pos"]), _dp(out["vel"])))
        return out

    def state(self, cloud=0):
        n = self.sizes[cloud]
        pos = np.zeros((n, 3)); vel = np.zeros((n, 3))
        _check(lib().sbk_get_state(self._h, cloud, _dp(pos), _dp(vel)))
        return pos, vel

    def status(self, cloud=0):
        st = CloudStatus()
        _check(lib().sbk_get_status(self._h, cloud, ctypes.byref(st)))
        return st.as_dict()

    def stats(self):
        st = Stats()
        _check(lib().sbk_get_stats(self._h, ctypes.byref(st)))
        return st.as_dict()

    def reset_stats(self):
        _check(lib().sbk_reset_stats(self._h))

    def eval_force(self, cloud=0):
        out = np.zeros((self.sizes[cloud], 3))
        _check(lib().sbk_eval_force(self._h, cloud, _dp(out)))
        return out

    def eval_coulomb(self, cloud=0, precision="fp64_exact"):
        p = PRECISIONS[precision] if isinstance(precision, str) else int(precision)
        out = np.zeros((self.sizes[cloud], 3))
        _check(lib().sbk_eval_coulomb(self._h, cloud, p, _dp(out)))
        return out

    def time_coulomb(self, precision="mixed", reps=5):
        p = PRECISIONS[precision] if isinstance(precision, str) else int(precision)
        ms = ctypes.c_double()
        _check(lib().sbk_time_coulomb(self._h, p, int(reps), ctypes.byref(ms)))
        return ms.value

    def load_tuning(self, table_path, precision="mixed"):
        """Applies the best shape recorded in an autotune table; False if none matches."""
        p = PRECISIONS[precision] if isinstance(precision, str) else int(precision)
        rc = lib().sbk_load_tuning(self._h, p, table_path.encode())
        if rc == -3:
            return False
        _check(rc)
        return True

    def autotune(self, precision="mixed", table_path=None):
        """Tries the launch shapes and Coulomb kernels and keeps the fastest:
        (targets/thread, source groups, slices, local coords, kernel)."""
        p = PRECISIONS[precision] if isinstance(precision, str) else int(precision)
        t = ctypes.c_int(); q = ctypes.c_int(); s = ctypes.c_int(); lc = ctypes.c_int()
        k = ctypes.c_int()
        _check(lib().sbk_autotune(self._h, p, ctypes.byref(t), ctypes.byref(q), ctypes.byref(s),
                                  ctypes.byref(lc), ctypes.byref(k),
                                  table_path.encode() if table_path else None))
        return t.value, q.value, s.value, lc.value, k.value

    def launch(self, precision="mixed"):
        """The launch the engine runs for `precision`, automatic choices resolved:
        (targets/thread, source groups, largest slice count, local coords, kernel)."""
        p = PRECISIONS[precision] if isinstance(precision, str) else int(precision)
        v = [ctypes.c_int() for _ in range(5)]
        _check(lib().sbk_get_launch(self._h, p, *[ctypes.byref(x) for x in v]))
        return tuple(x.value for x in v)

    def set_shape(self, shape):
        """Applies (targets/thread, source groups, slices, local coords[, kernel])."""
        self.set_launch(*shape[:4])
        if len(shape) > 4:
            self.set_coulomb_kernel(int(shape[4]))


def nbody_accel(pos):
    """The reference's NBODY call sequence (bin/nbody.cpp:24-76,216-271):
    initialize, simulation_Async on FP32 float4 positions, wait, finalize.
    Returns the FP32 accelerations sum_j (x_j-x_i)(|r|^2+1e-7)^-1.5."""
    import ctypes as C
    L = nbody_lib()
    p = np.zeros((len(pos), 4), dtype=np.float32)
    p[:, :3] = np.asarray(pos, dtype=np.float32).reshape(-1, 3)
    p[:, 3] = 1.0                                     # w = 1 "charge" (bin/nbody.cpp:121)
    res = np.zeros((len(pos), 3), dtype=np.float32)
    L.initialize(len(pos), False)
    try:
        L.simulation_Async(p.ctypes.data_as(C.c_void_p), res.ctypes.data_as(C.c_void_p),
                           len(pos), 0, 0, 0)
        L.simulation_Async_end(None, None, 0, 0, 0, 0)
    finally:
        L.finalize()
    return res


def cunbody1_force(xj, mj, xi, eps2, precision=None):
    """The reference's cunbody1_force (bin/cunbody.h:11-17) through the C ABI."""
    L = lib()
    xj = np.ascontiguousarray(xj, dtype=np.float64).reshape(-1, 3)
    xi = np.ascontiguousarray(xi, dtype=np.float64).reshape(-1, 3)
    mj = np.ascontiguousarray(mj, dtype=np.float64).ravel()
    ai = np.zeros_like(xi)
    old = None
    if precision is not None:
        p = PRECISIONS[precision] if isinstance(precision, str) else int(precision)
        old = L.sbk_cunbody_set_precision(p)
    try:
        L.cunbody1_force(_dp(xj), _dp(mj), _dp(xi), float(eps2), _dp(ai), xi.shape[0], xj.shape[0])
    finally:
        if old is not None:
            L.sbk_cunbody_set_precision(old)
    return ai


def cunbody_force_static(xj, mj, xi, eps2, precision=None):
    """cunbody1_force with no device allocation (sbk_cunbody_force_static):
    at most 32768 targets and sources.  Raises on an error."""
    L = lib()
    xj = np.ascontiguousarray(xj, dtype=np.float64).reshape(-1, 3)
    xi = np.ascontiguousarray(xi, dtype=np.float64).reshape(-1, 3)
    mj = np.ascontiguousarray(mj, dtype=np.float64).ravel()
    ai = np.zeros_like(xi)
    old = None
    if precision is not None:
        p = PRECISIONS[precision] if isinstance(precision, str) else int(precision)
        old = L.sbk_cunbody_set_precision(p)
    try:
        _check(L.sbk_cunbody_force_static(_dp(xj), _dp(mj), _dp(xi), float(eps2), _dp(ai),
                                          xi.shape[0], xj.shape[0]))
    finally:
        if old is not None:
            L.sbk_cunbody_set_precision(old)
    return ai
