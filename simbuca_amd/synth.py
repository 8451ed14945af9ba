"""Seeded synthetic ion clouds and the field set-ups of the benchmark workloads.

The reference initialises clouds from INI files (uniform ellipsoid, bin/
initialization.cpp:117-206); the workloads here need Gaussian positions and
Maxwell-Boltzmann velocities instead, so they are generated here:

* positions  ~ N(0, sigma) per axis,
* velocities ~ N(0, sqrt(kB T / m)) per axis,
* species drawn independently per ion with the given fractions,

all from ``numpy.random.Generator(PCG64(seed))`` in that order (species,
positions, velocities).  Masses are computed the way the reference does,
``(double)((long double)mass_amu * amu)`` (bin/ion.cpp:3-6).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

from . import AMU, EL_CHARGE, KB, KE, Fields, default_fields


def mass_kg(mass_amu):
    """Ion mass in kg with the reference's long-double product (bin/ion.cpp:3-6)."""
    m = np.asarray(mass_amu, dtype=np.float64)
    return (m.astype(np.longdouble) * np.longdouble(AMU)).astype(np.float64)


@dataclass
class Cloud:
    pos: np.ndarray
    vel: np.ndarray
    mass: np.ndarray
    charge: np.ndarray
    mass_amu: np.ndarray
    fields: Fields = None

    @property
    def n(self):
        return self.pos.shape[0]


def gaussian_cloud(n, sigma, temperature, species=((133.0, 1.0),), charge=1, seed=1):
    """n ions: Gaussian positions (sigma per axis) and Maxwell-Boltzmann
    velocities at `temperature`; `species` is a list of (mass_amu, fraction)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    masses = np.array([s[0] for s in species], dtype=np.float64)
    frac = np.array([s[1] for s in species], dtype=np.float64)
    frac = frac / frac.sum()
    if len(species) == 1:
        idx = np.zeros(n, dtype=np.int64)
    else:
        idx = np.searchsorted(np.cumsum(frac)[:-1], rng.random(n), side="right")
    m_amu = masses[idx]
    pos = rng.normal(0.0, sigma, (n, 3))
    m = mass_kg(m_amu)
    vel = rng.normal(0.0, 1.0, (n, 3)) * np.sqrt(KB * temperature / m)[:, None]
    q = np.full(n, charge, dtype=np.int32)
    return Cloud(pos=pos, vel=vel, mass=m, charge=q, mass_amu=m_amu)


def penning_frequencies(mass_amu, B=7.0, V0=10.0, d0=11e-3, charge=1):
    """(nu_c, nu_plus, nu_minus, nu_z) in Hz of an ideal Penning trap with the
    quadrupole V0/(2 d0^2)(z^2 - r^2/2)."""
    m = float(mass_kg(mass_amu))
    q = charge * EL_CHARGE
    wc = q * B / m
    wz2 = q * V0 / (m * d0 * d0)
    root = math.sqrt(wc * wc / 4.0 - wz2 / 2.0)
    wp, wm = wc / 2.0 + root, wc / 2.0 - root
    return wc / (2 * math.pi), wp / (2 * math.pi), wm / (2 * math.pi), math.sqrt(wz2) / (2 * math.pi)


def penning_fields(B=7.0, V0=10.0, d0=11e-3, coulomb=True, **kw) -> Fields:
    """Penning trap: uniform B along z and the ideal quadrupole; the fork's
    Paul-trap term (bin/force.cpp:70-75) is switched off."""
    f = default_fields(trap=0, Bz=B, V0=V0, d0=d0, coulomb=1 if coulomb else 0, coulomb_scale=1.0)
    for k, v in kw.items():
        setattr(f, k, v)
    return f


@dataclass
class Workload:
    name: str
    clouds: list
    t_sim: float
    rtol: float
    atol: float
    h0: float
    precision: str
    time_mode: int
    notes: dict = field(default_factory=dict)

    @property
    def n_total(self):
        return sum(c.n for c in self.clouds)


def workload(index: int, n_override=None) -> Workload:
    """The five benchmark set-ups (1-based index); `n_override` shrinks the
    ion count for tests."""
    from . import TIME_STAGE

    if index == 1:
        n = n_override or 500
        c = gaussian_cloud(n, 0.5e-3, 300.0, seed=1)
        c.fields = penning_fields()
        return Workload("500 ions 133u, B=7T V0=10V d0=11mm, DP5 rtol=1e-9, 100us", [c], 100e-6,
                        1e-9, 1e-9, 1e-9, "fp64_exact", TIME_STAGE, {"seed": 1})
    if index == 2:
        n = n_override or 10000
        c = gaussian_cloud(n, 1e-3, 300.0, species=((133.0, 0.5), (85.0, 0.5)), seed=2)
        nu_m = penning_frequencies(85.0)[2]
        c.fields = penning_fields(Ed=10.0, fd=nu_m)
        return Workload("10,000 ions 133u/85u, dipole 0.1 V/cm at nu-(85u), 50us", [c], 50e-6,
                        1e-9, 1e-9, 1e-9, "mixed", TIME_STAGE, {"seed": 2, "fd_Hz": nu_m})
    if index == 3:
        n = n_override or 65536
        c = gaussian_cloud(n, 1.5e-3, 300.0, seed=3)
        nu_c = penning_frequencies(133.0)[0]
        c.fields = penning_fields(Uq=0.1, rq=11e-3, fq=nu_c, delta=1e3)
        return Workload("65,536 ions 133u, quadrupole RF at nu_c + friction 1e3/s, 20us", [c],
                        20e-6, 1e-9, 1e-9, 1e-9, "mixed", TIME_STAGE,
                        {"seed": 3, "fq_Hz": nu_c, "Uq_V": 0.1, "rq_m": 11e-3})
    if index == 4:
        rng = np.random.Generator(np.random.PCG64(4))
        count = 256
        sizes = np.rint(np.exp(rng.uniform(math.log(64), math.log(4096), count))).astype(int)
        if n_override:
            sizes = np.minimum(sizes, n_override)
        nu_m = penning_frequencies(133.0)[2]
        clouds = []
        for k, nk in enumerate(sizes):
            c = gaussian_cloud(int(nk), 0.5e-3, 300.0, seed=1000 + k)
            fd = nu_m + (-5e3 + 10e3 * k / (count - 1))
            c.fields = penning_fields(Ed=10.0, fd=fd)
            clouds.append(c)
        return Workload("256 clouds 64..4096 ions, dipole sweep nu-(133u) +-5kHz, 100us", clouds,
                        100e-6, 1e-9, 1e-9, 1e-9, "mixed", TIME_STAGE,
                        {"seed": 4, "sizes_sum": int(sizes.sum())})
    if index == 5:
        n = n_override or 262144
        c = gaussian_cloud(n, 2e-3, 10.0, species=((133.0, 0.5), (85.0, 0.3), (39.0, 0.2)), seed=5)
        c.fields = penning_fields()
        return Workload("262,144 ions 133/85/39u, 10K, Kahan, 5us", [c], 5e-6, 1e-9, 1e-9, 1e-9,
                        "mixed_kahan", TIME_STAGE, {"seed": 5})
    raise ValueError(index)


def cloud_invariants(c: Cloud, pos, vel):
    """Total energy (kinetic + ideal Penning quadrupole + Coulomb) and
    canonical angular momentum L_z = sum m (x v_y - y v_x) + q e B (x^2+y^2)/2
    of cloud c at state (pos, vel), with the cloud's B and V0/d0.  Both are
    conserved for an axisymmetric trap without drive or friction."""
    B, V0, d0 = c.fields.Bz, c.fields.V0, c.fields.d0
    cpen = V0 / (2 * d0 * d0) if V0 else 0.0
    q = c.charge.astype(np.float64) * EL_CHARGE
    kin = 0.5 * (c.mass[:, None] * vel * vel).sum()
    trap = (q * cpen * (pos[:, 2] ** 2 - 0.5 * (pos[:, 0] ** 2 + pos[:, 1] ** 2))).sum()
    coul = 0.0
    for i in range(0, c.n, 512):
        d = pos[i:i + 512, None, :] - pos[None, :, :]
        r = np.sqrt((d * d).sum(-1))
        idx = np.arange(i, min(i + 512, c.n))
        r[idx - i, idx] = np.inf
        coul += 0.5 * KE * (q[i:i + 512, None] * q[None, :] / r).sum()
    lz = (c.mass * (pos[:, 0] * vel[:, 1] - pos[:, 1] * vel[:, 0])).sum() + \
        (q * B * (pos[:, 0] ** 2 + pos[:, 1] ** 2) / 2).sum()
    return kin + trap + coul, lz
