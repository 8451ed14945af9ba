"""The oracle's Penning excitation menu and field extras, pinned by what the
physics fixes.  The reference fork has none of these terms, so nothing of
the reference pins them.  Each term is checked against its potential (E =
-grad phi by central differences), against the symmetry it must have, and
against a closed-form trajectory integrated with the oracle's own DP5:

* octupole drive: E = -grad phi, the field is divergence-free, and it flips
  sign under a 45 degree turn;
* rotating wall, order 1 and 2: E = -grad phi, and the field turns with the
  phase (E_psi(R p) = R E_{psi - alpha}(p));
* axial dipole drive: the driven free ion z = A (1 - cos wt)/w^2;
* frequency sweeps: the driven free ion's velocity is a Fresnel integral;
* magnetic bottle: the field recovered from the Lorentz term is divergence-
  and curl-free, does no work, shifts the cyclotron frequency to
  q (B0 + B2 z0^2)/m on the plane z = z0, and pulls the orbit towards the
  centre with the mirror force -m v_perp^2 B2 z0 / B;
* buffer gas by mobility: v0 exp(-(q e/m) t / K) with K = K0 (p0/p)(T/T0),
  for two charge-to-mass ratios.

A plausible slip (a dropped factor 3, a sign, the 1/2 of the bottle's radial
term, a sweep phase without the 1/2) fails at least one of them.
"""
import math

import numpy as np
import pytest
from scipy.special import fresnel

import oracle as O

E = 1.60217733e-19


def _single(pos, vel, amu=133.0, charge=1):
    return (np.array([pos], dtype=np.float64), np.array([vel], dtype=np.float64),
            O.mass_kg([amu]), np.array([charge], dtype=np.int32))


def _field_at(points, t=0.0, **kw):
    """Electric field of the position-dependent terms: acceleration / (q e / m)."""
    n = len(points)
    m = O.mass_kg(np.full(n, 133.0))
    ch = np.ones(n, dtype=np.int32)
    f = O.fields(trap=0, **kw)
    a = O.force(np.asarray(points, dtype=np.float64), np.zeros((n, 3)), m, ch, f, t)
    return a / (E / m)[:, None]


def _grad(phi, p, h=None):
    # central differences with a step relative to the distance from the axis
    h = 1e-5 * math.hypot(p[0], p[1]) if h is None else h
    g = np.zeros(3)
    for k in range(3):
        d = np.zeros(3); d[k] = h
        g[k] = (phi(p + d) - phi(p - d)) / (2 * h)
    return g


def _rot(alpha):
    c, s = math.cos(alpha), math.sin(alpha)
    return np.array([[c, -s, 0.0], [s, c, 0.0], [0.0, 0.0, 1.0]])


# ---------------------------------------------------------------------------
# octupole drive
# ---------------------------------------------------------------------------

def test_octupole_is_minus_grad_of_its_potential(rng):
    Uo, ro, fo, phio, t = 3.0, 8e-3, 2.0e6, 0.4, 1.3e-7
    c = math.cos(2 * math.pi * fo * t + phio)

    def phi(p):
        u = complex(p[0], p[1])
        return -(Uo / (4 * ro ** 4)) * c * (u ** 4).real

    pts = rng.normal(0, 1e-3, (20, 3))
    Ef = _field_at(pts, t, Uo=Uo, ro=ro, fo=fo, phio=phio)
    for p, e in zip(pts, Ef):
        g = _grad(phi, p)
        assert np.allclose(e, -g, rtol=1e-6, atol=1e-9 * np.abs(g).max())
    # divergence-free: sum of d E_k / d x_k by central differences
    # (|E| = (Uo/ro^4) |c| rho^3 and its derivatives ~ |E|/rho, with rho the
    # distance from the axis)
    for p in pts[:5]:
        rho = math.hypot(p[0], p[1])
        div, h = 0.0, 1e-5 * rho
        for k in range(3):
            d = np.zeros(3); d[k] = h
            div += (_field_at([p + d], t, Uo=Uo, ro=ro, fo=fo, phio=phio)[0, k]
                    - _field_at([p - d], t, Uo=Uo, ro=ro, fo=fo, phio=phio)[0, k]) / (2 * h)
        scale = np.linalg.norm(_field_at([p], t, Uo=Uo, ro=ro, fo=fo, phio=phio)[0]) / rho
        assert abs(div) < 1e-7 * scale
    # order-4 pattern: a 45 degree turn flips the field
    R = _rot(math.pi / 4)
    for p, e in zip(pts[:5], Ef[:5]):
        e45 = _field_at([R @ p], t, Uo=Uo, ro=ro, fo=fo, phio=phio)[0]
        assert np.allclose(e45, -(R @ e), rtol=1e-9, atol=1e-12 * np.abs(e).max())


# ---------------------------------------------------------------------------
# rotating wall
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("order", [1, 2])
def test_rotating_wall_potential_and_rotation(rng, order):
    Uw, rw, fw, phiw, t = 0.7, 5e-3, 1.5e5, -0.3, 2.2e-6
    psi = 2 * math.pi * fw * t + phiw

    def phi(p):
        u = complex(p[0], p[1])
        return -(Uw / (order * rw ** order)) * (u ** order * complex(math.cos(order * psi),
                                                                    -math.sin(order * psi))).real

    kw = dict(Uw=Uw, rw=rw, fw=fw, phiw=phiw, rw_order=order)
    pts = rng.normal(0, 1e-3, (20, 3))
    Ef = _field_at(pts, t, **kw)
    for p, e in zip(pts, Ef):
        g = _grad(phi, p)
        assert np.allclose(e, -g, rtol=1e-6, atol=1e-9 * np.abs(g).max())
    if order == 1:
        assert np.allclose(np.hypot(Ef[:, 0], Ef[:, 1]), Uw / rw, rtol=1e-14)
    # the pattern turns with the phase: E_psi(R p) = R E_{psi - alpha}(p)
    alpha = 0.77
    R = _rot(alpha)
    back = _field_at(pts, t, **dict(kw, phiw=phiw - alpha))
    for p, e_back in zip(pts, back):
        e_rot = _field_at([R @ p], t, **kw)[0]
        assert np.allclose(e_rot, R @ e_back, rtol=1e-9, atol=1e-12 * np.abs(e_rot).max())


# ---------------------------------------------------------------------------
# axial drive and frequency sweeps (free ions, stage time)
# ---------------------------------------------------------------------------

def test_axial_drive_closed_form():
    p, v, m, ch = _single([0, 0, 0], [0, 0, 0])
    w = 2 * math.pi * 2e5
    f = O.fields(trap=0, Eax=5.0, fax=2e5)
    A = E * 5.0 / m[0]
    c = O.Cloud(p, v, m, ch, f, h0=1e-8, rtol=1e-12, atol=1e-16, time_mode=O.TIME_STAGE)
    c.advance(2.1e-5)
    pos, vel, sc = c.state()
    t = sc["lifetime"]
    assert pos[0, 2] == pytest.approx(A * (1 - math.cos(w * t)) / w ** 2, rel=1e-8)
    assert vel[0, 2] == pytest.approx(A * math.sin(w * t) / w, rel=1e-8, abs=1e-9 * A / w)
    assert pos[0, 0] == 0.0 and pos[0, 1] == 0.0


def _fresnel_integral(f, r, t):
    """int_0^t cos(2 pi (f tau + r tau^2 / 2)) dtau by completing the square."""
    s = math.sqrt(2 * r)
    b = math.pi * f * f / r
    u0, u1 = (f / r) * s, (t + f / r) * s
    S1, C1 = fresnel(u1)
    S0, C0 = fresnel(u0)
    return (math.cos(b) * (C1 - C0) + math.sin(b) * (S1 - S0)) / s


@pytest.mark.parametrize("which", ["dipole", "quadrupole"])
def test_swept_drive_closed_form(which):
    # sweep 0 -> 2 MHz over 20 us: the phase must be 2 pi (f t + r t^2 / 2)
    f0, r, T = 1e5, 1e11, 2e-5
    m = O.mass_kg([133.0])
    if which == "dipole":
        p, v, m, ch = _single([0, 0, 0], [0, 0, 0])
        f = O.fields(trap=0, Ed=8.0, fd=f0, fd_rate=r)
        A = E * 8.0 / m[0]
        k = 0
    else:
        # the quadrupole drive is linear in x: on the line y = 0 a free ion held
        # at fixed x would feel (Uq/rq^2) x cos(psi); use a tiny x so the ion
        # barely moves and compare the velocity with the fixed-x integral
        x0 = 1e-3
        p, v, m, ch = _single([x0, 0, 0], [0, 0, 0])
        f = O.fields(trap=0, Uq=1e-6, rq=1e-2, fq=f0, fq_rate=r)
        A = E * (1e-6 / 1e-4) * x0 / m[0]
        k = 0
    c = O.Cloud(p, v, m, ch, f, h0=1e-9, rtol=1e-12, atol=1e-18, time_mode=O.TIME_STAGE)
    c.advance(T)
    _, vel, sc = c.state()
    t = sc["lifetime"]
    want = A * _fresnel_integral(f0, r, t)
    assert vel[0, k] == pytest.approx(want, rel=1e-6, abs=1e-9 * A * T)
    # without the 1/2 of the sweep phase the answer would differ by far more
    wrong = A * _fresnel_integral(f0, 2 * r, t)
    assert abs(wrong - want) > 1e-2 * abs(want)


# ---------------------------------------------------------------------------
# magnetic bottle
# ---------------------------------------------------------------------------

def _bottle_B(points, B2):
    """B recovered from the Lorentz term: v = x_hat gives q/m (0, -Bz, By),
    v = y_hat gives q/m (Bz, 0, -Bx)."""
    n = len(points)
    m = O.mass_kg(np.full(n, 133.0)); ch = np.ones(n, dtype=np.int32)
    f = O.fields(trap=0, B2=B2)
    P = np.asarray(points, dtype=np.float64)
    ax = O.force(P, np.tile([1.0, 0, 0], (n, 1)), m, ch, f, 0.0) / (E / m)[:, None]
    ay = O.force(P, np.tile([0, 1.0, 0], (n, 1)), m, ch, f, 0.0) / (E / m)[:, None]
    return np.stack([-ay[:, 2], ax[:, 2], ay[:, 0]], axis=1), ax, ay


def test_magnetic_bottle_is_a_vacuum_field(rng):
    B2 = 3.0e4
    pts = rng.normal(0, 2e-3, (12, 3))
    B, ax, ay = _bottle_B(pts, B2)
    assert np.allclose(-ax[:, 1], B[:, 2], rtol=1e-14, atol=0)     # the two probes agree
    h = 1e-6
    for p in pts:
        J = np.zeros((3, 3))                                        # J[i, k] = d B_i / d x_k
        for k in range(3):
            d = np.zeros(3); d[k] = h
            J[:, k] = (_bottle_B([p + d], B2)[0][0] - _bottle_B([p - d], B2)[0][0]) / (2 * h)
        scale = B2 * np.abs(p).max()
        assert abs(np.trace(J)) < 1e-6 * scale                      # div B = 0
        assert np.abs(J - J.T).max() < 1e-6 * scale                  # curl B = 0
    # B2 = 0 at the origin, and the axial field grows as B2 z^2 on the axis
    Bz_axis = _bottle_B([[0, 0, 1e-3]], B2)[0][0]
    assert Bz_axis[2] == pytest.approx(B2 * 1e-6, rel=1e-14) and Bz_axis[0] == 0.0


def test_magnetic_bottle_does_no_work_and_shifts_the_cyclotron_frequency(rng):
    B0, B2, z0 = 5.0, 2.0e5, 2e-3
    pos = rng.normal(0, 1e-3, (16, 3)); vel = rng.normal(0, 300.0, (16, 3))
    m = O.mass_kg(np.full(16, 133.0)); ch = np.ones(16, dtype=np.int32)
    a = O.force(pos, vel, m, ch, O.fields(trap=0, Bz=B0, B2=B2), 0.0)
    assert np.abs((a * vel).sum(1)).max() < 1e-13 * (np.abs(a) * np.abs(vel)).sum(1).max()
    # a small cyclotron orbit around the axis on the plane z = z0
    m1 = O.mass_kg([133.0])[0]
    w = E * (B0 + B2 * z0 ** 2) / m1
    r = 1e-7
    p, v, mm, chh = _single([r, 0, z0], [0, -w * r, 0])
    c = O.Cloud(p, v, mm, chh, O.fields(trap=0, Bz=B0, B2=B2), h0=1e-9, rtol=1e-12, atol=1e-18,
                time_mode=O.TIME_STAGE)
    c.advance(4e-6)
    P, _, sc = c.state()
    t = sc["lifetime"]
    assert P[0, 0] == pytest.approx(r * math.cos(w * t), abs=1e-4 * r)
    assert P[0, 1] == pytest.approx(-r * math.sin(w * t), abs=1e-4 * r)
    # the mirror force of the radial field, q (v x B)_z = -m v_perp^2 B2 z / B,
    # pulls the orbit towards the bottle's centre: z - z0 = -(v_perp^2 B2 z0 / B) t^2 / 2
    drift = -0.5 * (w * r) ** 2 * B2 * z0 / (B0 + B2 * z0 ** 2) * t * t
    assert P[0, 2] - z0 == pytest.approx(drift, rel=1e-3)


# ---------------------------------------------------------------------------
# buffer gas by mobility
# ---------------------------------------------------------------------------

def test_buffer_gas_mobility_closed_form():
    K0, p_mbar, T_gas = 1.6e-4, 2e-3, 300.0
    K = K0 * (1013.25 / p_mbar) * (T_gas / 273.15)
    pos = np.zeros((2, 3)); vel = np.array([[300.0, -40.0, 10.0], [-50.0, 80.0, 0.0]])
    m = O.mass_kg([133.0, 39.0]); ch = np.array([1, 2], dtype=np.int32)
    f = O.fields(trap=0, K0=K0, p_mbar=p_mbar, T_gas=T_gas)
    c = O.Cloud(pos, vel, m, ch, f, h0=1e-7, rtol=1e-12, atol=1e-16, time_mode=O.TIME_STAGE)
    c.advance(5e-4)
    P, V, sc = c.state()
    t = sc["lifetime"]
    for i in range(2):
        g = ch[i] * E / m[i] / K
        assert np.allclose(V[i], vel[i] * math.exp(-g * t), rtol=1e-8)
        assert np.allclose(P[i], vel[i] * (1 - math.exp(-g * t)) / g, rtol=1e-8)
    # at normal conditions the mobility is K0 itself
    f0 = O.fields(trap=0, K0=K0, p_mbar=1013.25, T_gas=273.15)
    a = O.force(pos, vel, m, ch, f0, 0.0)
    assert np.allclose(a, -(ch * E / m / K0)[:, None] * vel, rtol=1e-14)


# ---------------------------------------------------------------------------
# SWIFT: sampled waveform as a dipole or a quadrupole
# ---------------------------------------------------------------------------

def test_swift_constant_waveform_is_a_static_field(rng):
    # a constant U gives E = U/r along x: the TOF kick's form (bin/force.cpp:77-80)
    pos = rng.normal(0, 1e-3, (9, 3)); vel = rng.normal(0, 100.0, (9, 3))
    m = O.mass_kg(np.full(9, 133.0)); ch = np.ones(9, dtype=np.int32)
    f = O.set_swift(O.fields(trap=0), np.full(17, 2.0), 1e-6, 1e-7, 4e-3, kind=1)
    a = O.force(pos, vel, m, ch, f, 1.73e-6)
    ref = O.force(pos, vel, m, ch, O.fields(trap=0, tof=1, E_kick=2.0 / 4e-3), 1.73e-6)
    assert np.array_equal(a, ref)
    # outside the sampled window there is no field at all
    assert np.array_equal(O.force(pos, vel, m, ch, f, 0.9e-6), np.zeros((9, 3)))
    assert np.array_equal(O.force(pos, vel, m, ch, f, 1e-6 + 16.5e-7), np.zeros((9, 3)))


def test_swift_interpolation_and_quadrupole_form(rng):
    # a sampled cosine equals the quadrupole drive to the interpolation error
    # (w dt)^2 / 8, and a linear waveform is interpolated exactly
    Uq, rq, fq = 0.4, 9e-3, 3e5
    dt = 1e-8
    t = np.arange(0, 401) * dt
    f = O.set_swift(O.fields(trap=0), Uq * np.cos(2 * math.pi * fq * t), 0.0, dt, rq, kind=2)
    g = O.fields(trap=0, Uq=Uq, rq=rq, fq=fq)
    pos = rng.normal(0, 1e-3, (6, 3)); vel = np.zeros((6, 3))
    m = O.mass_kg(np.full(6, 85.0)); ch = np.ones(6, dtype=np.int32)
    # |U_interp - U| <= (w dt)^2 / 8 * Uq for a linear interpolant of Uq cos(w t)
    amp = (E / m[0]) * Uq / rq ** 2 * np.abs(pos[:, :2]).max()
    bound = (2 * math.pi * fq * dt) ** 2 / 8 * amp
    for tt in rng.uniform(0, 400 * dt, 7):
        a = O.force(pos, vel, m, ch, f, tt)
        b = O.force(pos, vel, m, ch, g, tt)
        assert np.abs(a - b).max() <= 1.01 * bound
        assert np.all(a[:, 2] == 0.0)
        # and the field has the quadrupole's form: a_x / x = -a_y / y
        assert np.allclose(a[:, 0] / pos[:, 0], -a[:, 1] / pos[:, 1], rtol=1e-12)
    lin = O.set_swift(O.fields(trap=0), 3.0 + 2.0 * np.arange(5), 0.0, 1.0, 1.0, kind=1)
    E_at = lambda tt: O.force(np.zeros((1, 3)), np.zeros((1, 3)), m[:1], ch[:1], lin, tt)[0, 0] / (E / m[0])
    for tt in (0.0, 0.25, 1.5, 2.75, 4.0):
        assert E_at(tt) == pytest.approx(3.0 + 2.0 * tt, rel=1e-15)


def test_swift_driven_free_ion():
    # a sampled cos(w t) drive: the free ion follows A (1 - cos w t) / w^2 up to
    # the interpolation error of the drive
    p, v, m, ch = _single([0, 0, 0], [0, 0, 0])
    w = 2 * math.pi * 1e5
    dt = 2e-9
    t = np.arange(0, 20001) * dt
    f = O.set_swift(O.fields(trap=0), 6.0 * np.cos(w * t), 0.0, dt, 1.0, kind=1)
    A = E * 6.0 / m[0]
    c = O.Cloud(p, v, m, ch, f, h0=1e-8, rtol=1e-11, atol=1e-16, time_mode=O.TIME_STAGE)
    c.advance(3.3e-5)
    pos, vel, sc = c.state()
    tt = sc["lifetime"]
    assert pos[0, 0] == pytest.approx(A * (1 - math.cos(w * tt)) / w ** 2, rel=1e-5)


# ---------------------------------------------------------------------------
# non-ideal trap (C4, C6) and axisymmetric field maps
# ---------------------------------------------------------------------------
V0, D0 = 10.0, 11e-3


def _legendre_phi(k, Ck):
    """(V0/2) C_k (rho/d0)^k P_k(cos theta) from scipy's Legendre polynomials:
    an independent route to the potential whose gradient the oracle uses."""
    from scipy.special import eval_legendre

    def phi(p):
        rho = float(np.linalg.norm(p))
        return 0.5 * V0 * Ck * (rho / D0) ** k * eval_legendre(k, p[2] / rho)
    return phi


@pytest.mark.parametrize("k,Ck", [(4, 0.03), (6, -0.02)])
def test_anharmonic_terms_are_minus_grad_of_the_legendre_potential(rng, k, Ck):
    kw = {"C4": Ck} if k == 4 else {"C6": Ck}
    pts = rng.uniform(-4e-3, 4e-3, (12, 3))
    E_all = _field_at(pts, V0=V0, d0=D0, **kw)
    E_ideal = _field_at(pts, V0=V0, d0=D0)
    phi = _legendre_phi(k, Ck)
    for p, e in zip(pts, E_all - E_ideal):
        want = -_grad(phi, p, h=3e-7)
        # central differences err by about (h/rho)^2; a slipped coefficient errs by O(1)
        assert np.allclose(e, want, rtol=1e-5, atol=1e-6 * np.abs(want).max()), (k, p, e, want)


@pytest.mark.parametrize("kw", [{"C4": 0.05}, {"C6": 0.05}, {"C4": -0.01, "C6": 0.02}])
def test_anharmonic_field_is_divergence_free(rng, kw):
    # a potential field in vacuum: div E = 0 (the Laplace condition that fixes
    # the radial coefficients once the axial ones are given)
    h = 1e-6
    for p in rng.uniform(-4e-3, 4e-3, (8, 3)):
        div, scale = 0.0, 0.0
        for k in range(3):
            d = np.zeros(3); d[k] = h
            dk = (_field_at([p + d], V0=V0, d0=D0, **kw)[0, k]
                  - _field_at([p - d], V0=V0, d0=D0, **kw)[0, k]) / (2 * h)
            div += dk
            scale = max(scale, abs(dk))
        assert abs(div) <= 1e-6 * scale, (p, div, scale)


def _axial_frequency(C4=0.0, C6=0.0, A=2e-3, periods=10, per_period=400):
    """Angular frequency of an ion released at rest at z = A on the axis,
    from the zero crossings of z(t) integrated by the oracle's DP5."""
    m = O.mass_kg([133.0])[0]
    w0 = math.sqrt(E * V0 / (m * D0 * D0))
    h = 2 * math.pi / w0 / per_period
    p, v, mm, ch = _single([0.0, 0.0, A], [0.0, 0.0, 0.0])
    c = O.Cloud(p, v, mm, ch, O.fields(trap=0, V0=V0, d0=D0, C4=C4, C6=C6), h0=h, adaptive=False)
    ts, zs = [0.0], [A]
    for _ in range(periods * per_period):
        c.step()
        pos, _, sc = c.state()
        ts.append(sc["lifetime"]); zs.append(pos[0, 2])
    ts, zs = np.array(ts), np.array(zs)
    k = np.nonzero(np.sign(zs[:-1]) != np.sign(zs[1:]))[0]
    tc = ts[k] - zs[k] * (ts[k + 1] - ts[k]) / (zs[k + 1] - zs[k])
    half = np.diff(tc).mean()
    return math.pi / half, w0


@pytest.mark.parametrize("C4,C6,A,want", [
    (0.02, 0.0, 2e-3, lambda A: 0.75 * 0.02 * (A / D0) ** 2),
    (0.0, 0.1, 3e-3, lambda A: 15.0 / 16.0 * 0.1 * (A / D0) ** 4),
    (-0.03, 0.0, 2.5e-3, lambda A: 0.75 * -0.03 * (A / D0) ** 2),
])
def test_axial_frequency_shifts_with_amplitude(C4, C6, A, want):
    # on the axis z'' = -w0^2 (z + 2 C4 z^3/d0^2 + 3 C6 z^5/d0^4); harmonic
    # balance gives dw/w0 = (3/4) C4 (A/d0)^2 and (15/16) C6 (A/d0)^4 to first order
    w, w0 = _axial_frequency(C4, C6, A)
    w_ideal, _ = _axial_frequency(0.0, 0.0, A)
    assert abs(w_ideal / w0 - 1.0) < 1e-9                # the ideal trap: exactly w0
    shift = w / w0 - 1.0
    assert shift == pytest.approx(want(A), rel=0.02), (shift, want(A))


def test_anharmonic_radial_force_in_the_midplane():
    # z = 0: phi4 = c4k (3/8) r^4 and phi6 = -c6k (5/16) r^6, so the radial field is
    # -(3/2) c4k r^3 and +(15/8) c6k r^5
    r = 3e-3
    c4k = V0 * 0.04 / (2 * D0 ** 4)
    c6k = V0 * 0.07 / (2 * D0 ** 6)
    e4 = _field_at([[r, 0.0, 0.0]], V0=V0, d0=D0, C4=0.04)[0] - _field_at([[r, 0.0, 0.0]], V0=V0, d0=D0)[0]
    e6 = _field_at([[0.0, r, 0.0]], V0=V0, d0=D0, C6=0.07)[0] - _field_at([[0.0, r, 0.0]], V0=V0, d0=D0)[0]
    assert e4[0] == pytest.approx(-1.5 * c4k * r ** 3, rel=1e-9) and abs(e4[2]) == 0.0
    assert e6[1] == pytest.approx(1.875 * c6k * r ** 5, rel=1e-9) and abs(e6[2]) == 0.0


def _grid(nr, nz, dr, z0, dz, field):
    r = np.arange(nr) * dr
    z = z0 + np.arange(nz) * dz
    R, Z = np.meshgrid(r, z, indexing="ij")
    er, ez = field(R, Z)
    return er, ez


def test_field_map_of_a_linear_field_is_exact(rng):
    # the ideal quadrupole E = cpen (r, -2 z) is linear in r and in z, which
    # bilinear interpolation reproduces: the map equals the V0 term
    cpen = V0 / (2 * D0 * D0)
    er, ez = _grid(41, 61, 2.5e-4, -7.5e-3, 2.5e-4, lambda R, Z: (cpen * R, -2.0 * cpen * Z))
    f = O.set_emap(O.fields(trap=0), er, ez, 2.5e-4, -7.5e-3, 2.5e-4)
    pts = np.column_stack([rng.uniform(-6e-3, 6e-3, (50, 2)), rng.uniform(-7e-3, 7e-3, 50)])
    pts = pts[np.hypot(pts[:, 0], pts[:, 1]) < 9.5e-3]
    m = O.mass_kg(np.full(len(pts), 133.0)); ch = np.ones(len(pts), np.int32)
    a_map = O.force(pts, np.zeros_like(pts), m, ch, f, 0.0)
    a_v0 = O.force(pts, np.zeros_like(pts), m, ch, O.fields(trap=0, V0=V0, d0=D0), 0.0)
    assert np.allclose(a_map, a_v0, rtol=1e-11, atol=1e-11 * np.abs(a_v0).max())
    # scale multiplies; outside the grid the map adds nothing; on the axis no radial part
    f2 = O.set_emap(O.fields(trap=0), er, ez, 2.5e-4, -7.5e-3, 2.5e-4, scale=-2.0)
    assert np.allclose(O.force(pts, np.zeros_like(pts), m, ch, f2, 0.0), -2.0 * a_map, rtol=1e-14)
    out = np.array([[0.0, 0.0, 8e-3], [1.2e-2, 0.0, 0.0], [0.0, -1.1e-2, 1e-3]])
    m3 = O.mass_kg(np.full(3, 133.0)); c3 = np.ones(3, np.int32)
    assert np.array_equal(O.force(out, np.zeros((3, 3)), m3, c3, f, 0.0), np.zeros((3, 3)))
    axis = np.array([[0.0, 0.0, 1e-3]])
    a_axis = O.force(axis, np.zeros((1, 3)), O.mass_kg([133.0]), np.ones(1, np.int32), f, 0.0)
    assert a_axis[0, 0] == 0.0 and a_axis[0, 1] == 0.0 and np.isfinite(a_axis).all()


def test_field_map_converges_at_second_order():
    # the C4 field (nonlinear) sampled on grids of spacing h and h/2: the
    # interpolation error falls about fourfold
    c4k = V0 * 0.05 / (2 * D0 ** 4)

    def c4_field(R, Z):   # E = -grad phi4 in (r, z)
        return c4k * R * (6.0 * Z * Z - 1.5 * R * R), c4k * Z * (6.0 * R * R - 4.0 * Z * Z)

    g = np.random.Generator(np.random.PCG64(17))
    pts = np.column_stack([g.uniform(-3e-3, 3e-3, (3000, 2)), g.uniform(-4e-3, 4e-3, 3000)])
    m = O.mass_kg(np.full(len(pts), 133.0)); ch = np.ones(len(pts), np.int32)
    exact = (O.force(pts, np.zeros_like(pts), m, ch, O.fields(trap=0, V0=V0, d0=D0, C4=0.05), 0.0)
             - O.force(pts, np.zeros_like(pts), m, ch, O.fields(trap=0, V0=V0, d0=D0), 0.0))
    errs = []
    for h in (2e-4, 1e-4, 5e-5):
        nr, nz = int(round(5e-3 / h)) + 1, int(round(1e-2 / h)) + 1
        er, ez = _grid(nr, nz, h, -5e-3, h, c4_field)
        f = O.set_emap(O.fields(trap=0), er, ez, h, -5e-3, h)
        a = O.force(pts, np.zeros_like(pts), m, ch, f, 0.0)
        errs.append(np.sqrt(((a - exact) ** 2).mean() / (exact ** 2).mean()))   # rms
    assert errs[0] / errs[1] > 3.6 and errs[1] / errs[2] > 3.6, errs
