"""The CPU oracle pinned against the reference's golden vectors, the live
reference build (when present) and closed forms of the physics.

Each closed-form test is chosen so that a plausible mistake in the oracle
(dropped charge factor, wrong sign, wrong tableau entry, wrong time) fails.
"""
import math

import numpy as np
import pytest

import oracle as O
from conftest import TRAP_KEYS, golden_fields, load_golden

E = 1.60217733e-19
KE = 8.987551787e9


def _fo(**kw):
    return O.fields(**kw)


# ---------------------------------------------------------------------------
# golden vectors produced by the reference's own force()/step()
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("name", ["force_n500_trap_tof_coulomb.npz", "force_n257_coulomb_scaled.npz"])
def test_force_matches_reference_golden(name):
    g = load_golden(name)
    f = golden_fields(g["fields"], _fo)
    m = O.mass_kg(g["mass_amu"])
    out = O.force(g["pos"], g["vel"], m, g["charge"], f, float(g["lifetime"]))
    assert np.array_equal(out, g["derivs"])   # bit-identical


@pytest.mark.parametrize("name", ["steps_adaptive_n64.npz", "steps_fixed_rf_n64.npz"])
def test_steps_match_reference_golden(name):
    g = load_golden(name)
    f = golden_fields(g["fields"], _fo)
    m = O.mass_kg(g["mass_amu"])
    c = O.Cloud(g["pos"], g["vel"], m, g["charge"], f, lifetime=float(g["lifetime"]),
                h0=float(g["h0"]), rtol=float(g["rtol"]), atol=float(g["atol"]),
                adaptive=bool(g["adaptive"]), time_mode=O.TIME_FROZEN)
    hist = []
    for _ in range(int(g["nsteps"])):
        hist.append(c.state()[2]["h"])
        c.step()
    pos, vel, sc = c.state()
    assert np.array_equal(pos, g["pos_out"])
    assert np.array_equal(vel, g["vel_out"])
    assert np.array_equal(np.array(hist), g["hist"][:, 0])
    assert sc["lifetime"] == g["scalars"][0]
    assert sc["h"] == g["scalars"][1]
    if g["adaptive"]:
        assert sc["attempts"] == g["scalars"][5]   # same accept/reject sequence


@pytest.mark.skipif(not O.ref_available(), reason="reference build oracle/_ref absent")
def test_force_matches_live_reference(rng):
    n = 300
    amu = rng.choice([133.0, 85.0, 39.0], n)
    ch = rng.choice(np.array([1, 2, 3, -1], dtype=np.int32), n)
    pos = rng.normal(0, 8e-4, (n, 3)); vel = rng.normal(0, 100, (n, 3))
    f = _fo(trap=1, Vrf=30.0, Vdc=-7.0, kappa=0.3, r0=0.01, z0=0.02, freq_rf=1.1e6, tof=1,
            E_kick=-2.5, coulomb=1, coulomb_scale=2.0)
    out = O.force(pos, vel, O.mass_kg(amu), ch, f, 7.7e-7)
    ref = O.ref_force(amu, ch, pos, vel, f, 7.7e-7)
    assert np.array_equal(out, ref)


# ---------------------------------------------------------------------------
# Coulomb sum: closed forms, invariants, brute force
# ---------------------------------------------------------------------------

def test_two_ions_closed_form():
    m = O.mass_kg([133.0, 85.0])
    ch = np.array([2, -1], dtype=np.int32)
    pos = np.array([[1e-4, -2e-4, 3e-4], [-3e-4, 1e-4, 0.5e-4]])
    f = _fo(trap=0, coulomb=1)
    a = O.force(pos, np.zeros((2, 3)), m, ch, f, 0.0)
    d = pos[0] - pos[1]
    r = np.linalg.norm(d)
    # a_0 = ke e^2 q0 q1 / m0 * d / r^3 (attractive here: q0 q1 < 0)
    want0 = KE * E * E * 2 * (-1) / m[0] * d / r ** 3
    want1 = -KE * E * E * 2 * (-1) / m[1] * d / r ** 3
    assert np.allclose(a[0], want0, rtol=1e-14, atol=0)
    assert np.allclose(a[1], want1, rtol=1e-14, atol=0)
    assert np.dot(a[0], d) < 0   # opposite charges attract


def test_newton_third_law(rng):
    n = 200
    m = O.mass_kg(rng.choice([133.0, 85.0], n))
    ch = rng.choice(np.array([1, 2, -1], dtype=np.int32), n)
    pos = rng.normal(0, 1e-3, (n, 3))
    a = O.force(pos, np.zeros((n, 3)), m, ch, _fo(trap=0, coulomb=1), 0.0)
    p = (m[:, None] * a).sum(0)
    scale = np.abs(m[:, None] * a).sum(0)
    assert np.all(np.abs(p) <= 1e-13 * scale)


def test_coulomb_brute_force_long_double(rng):
    n = 48
    m = O.mass_kg(rng.choice([133.0, 39.0], n))
    ch = rng.choice(np.array([1, 3, -2], dtype=np.int32), n)
    pos = rng.normal(0, 1e-3, (n, 3))
    cs = 0.75
    rows = np.arange(n, dtype=np.int32)
    f = O.coulomb_rows(pos, m, ch, cs, rows)
    P = pos.astype(np.longdouble)
    want = np.zeros((n, 3), dtype=np.longdouble)
    for j in range(n):
        d = P - P[j]
        r2 = (d * d).sum(1)
        r2[j] = 1
        w = ch.astype(np.longdouble) / r2 ** np.longdouble(1.5)
        w[j] = 0
        want[j] = (d * w[:, None]).sum(0) * np.longdouble(cs * KE * E * E) * ch[j] / m[j]
    err = np.abs(f - want.astype(np.float64)).max() / np.abs(want).max()
    assert err < 1e-14


def test_cunbody_interface_semantics(rng):
    n = 40
    x = rng.normal(0, 1e-3, (n, 3)); mj = rng.uniform(0.5, 2.0, n)
    eps2 = 1e-8
    ai = O.cunbody_force(x, mj, x, eps2)
    want = np.zeros((n, 3))
    for i in range(n):
        d = x - x[i]
        r2 = (d * d).sum(1) + eps2
        want[i] = (mj[:, None] * d / r2[:, None] ** 1.5).sum(0)
    assert np.allclose(ai, want, rtol=1e-13, atol=1e-13 * np.abs(want).max())


def test_cunbody_interface_reproduces_the_golden_force():
    """The reference's CUNBODY call (bin/cunbody.cpp:65-66: xj = pos2, mj = 1,
    xi = a copy of pos2) scaled as its caller does, derivs_j -=
    coulomb_scale e^2 ke / m_j * ai_j (bin/force.cpp:87-93), must give the
    reference's own CPU force() for unit charges without trap or TOF terms
    (force_n257_coulomb_scaled.npz, coulomb_scale 0.5).  Same summation order
    per target; only where the prefactor multiplies differs, so the two agree
    to a few ulps.  A dropped mass, a sign or a swapped xi/xj fails this."""
    g = load_golden("force_n257_coulomb_scaled.npz")
    assert (g["charge"] == 1).all()
    fields = dict(zip(TRAP_KEYS, g["fields"].tolist()))
    assert fields["trap"] == 0 and fields["tof"] == 0 and fields["coulomb"] == 1
    m = O.mass_kg(g["mass_amu"])
    x = np.ascontiguousarray(g["pos"])
    scale = fields["coulomb_scale"] * E * E * KE / m
    for eps2 in (0.0, float(np.float32(1e-23))):
        ai = O.cunbody_force(x, np.ones(len(x)), x.copy(), eps2)
        assert np.isfinite(ai).all()
        derivs = -scale[:, None] * ai
        err = np.linalg.norm(derivs - g["derivs"]) / np.linalg.norm(g["derivs"])
        assert err < (1e-14 if eps2 == 0.0 else 1e-10), (eps2, err)
    # a wrong sign or a missing mass fails
    assert np.linalg.norm(scale[:, None] * ai - g["derivs"]) > 0.5 * np.linalg.norm(g["derivs"])


def test_cunbody_interface_skips_the_self_pair_only_for_aliased_sets(rng):
    x = rng.normal(0, 1e-3, (17, 3))
    a0 = O.cunbody_force(x, np.ones(17), x.copy(), 0.0)
    assert np.isfinite(a0).all()
    # a different target set holding one of the sources: that pair is 0/0
    xi = x[:5].copy()
    a1 = O.cunbody_force(x, np.ones(17), xi, 0.0)
    assert np.isnan(a1).all(axis=1).all()
    a2 = O.cunbody_force(x, np.ones(17), xi, 1e-20)
    assert np.isfinite(a2).all()
    assert np.allclose(a2, a0[:5], rtol=1e-12, atol=1e-12 * np.abs(a0).max())


# ---------------------------------------------------------------------------
# trap and extension terms
# ---------------------------------------------------------------------------

def test_trap_quadrupole_is_laplacian_free_and_linear(rng):
    # a = qe/m (gx x, gy y, gz z) with gx + gy + gz = 0 (Laplace) at any t
    m = O.mass_kg([133.0]); ch = np.array([1], dtype=np.int32)
    f = _fo(trap=1, Vrf=40.0, Vdc=3.0, kappa=0.2, r0=0.01, z0=0.015, freq_rf=1e6)
    t = 3.3e-7
    g = []
    for axis in range(3):
        p = np.zeros((1, 3)); p[0, axis] = 1e-3
        g.append(O.force(p, np.zeros((1, 3)), m, ch, f, t)[0, axis] / 1e-3 / (E / m[0]))
    assert abs(sum(g)) < 1e-12 * max(abs(x) for x in g)
    c = math.cos(2 * math.pi * 1e6 * t)
    assert g[0] == pytest.approx(0.2 * 3.0 / 0.015 ** 2 - 40.0 * c / 0.01 ** 2, rel=1e-13)


def test_penning_quadrupole_equals_reference_dc_term():
    # V0/(2 d0^2) == kappa Vdc / z0^2 with the RF off
    m = O.mass_kg([133.0]); ch = np.array([1], dtype=np.int32)
    p = np.array([[3e-4, -2e-4, 5e-4]])
    fa = _fo(trap=0, V0=10.0, d0=11e-3)
    fb = _fo(trap=1, Vrf=0.0, kappa=1.0, Vdc=5.0, z0=11e-3, r0=1.0)
    a = O.force(p, np.zeros((1, 3)), m, ch, fa, 0.0)
    b = O.force(p, np.zeros((1, 3)), m, ch, fb, 0.0)
    assert np.allclose(a, b, rtol=1e-15, atol=0)


def test_quadrupole_excitation_equals_reference_rf_term():
    # (Uq/rq^2) cos(w t)(x,-y,0) is the fork's RF term with Vrf = -Uq
    m = O.mass_kg([85.0]); ch = np.array([2], dtype=np.int32)
    p = np.array([[3e-4, -2e-4, 0.0]])
    t = 1.7e-7
    fa = _fo(trap=0, Uq=0.3, rq=5e-3, fq=8e5)
    fb = _fo(trap=1, kappa=0.0, Vrf=-0.3, r0=5e-3, z0=1.0, freq_rf=8e5)
    a = O.force(p, np.zeros((1, 3)), m, ch, fa, t)
    b = O.force(p, np.zeros((1, 3)), m, ch, fb, t)
    assert np.allclose(a, b, rtol=1e-14, atol=0)


def test_lorentz_and_friction_terms():
    m = O.mass_kg([133.0]); ch = np.array([1], dtype=np.int32)
    v = np.array([[120.0, -40.0, 15.0]])
    f = _fo(trap=0, Bz=7.0, delta=1e3)
    a = O.force(np.zeros((1, 3)), v, m, ch, f, 0.0)[0]
    qm = E / m[0]
    want = qm * np.cross(v[0], [0, 0, 7.0]) - 1e3 * v[0]
    assert np.allclose(a, want, rtol=1e-14, atol=1e-12)


# ---------------------------------------------------------------------------
# Dormand-Prince 5(4), controller and step
# ---------------------------------------------------------------------------

def _single(pos, vel, amu=133.0, charge=1):
    return (np.array([pos], dtype=np.float64), np.array([vel], dtype=np.float64),
            O.mass_kg([amu]), np.array([charge], dtype=np.int32))


def test_dp5_exact_for_quadratic_motion():
    # constant TOF field: x(t) = x0 + v0 t + a t^2/2 is integrated exactly
    p, v, m, ch = _single([1e-4, 0, 0], [50.0, 0, 0])
    f = _fo(trap=0, tof=1, E_kick=2.0)
    h = 1e-7
    c = O.Cloud(p, v, m, ch, f, h0=h, adaptive=False)
    for _ in range(10):
        c.step()
    pos, vel, sc = c.state()
    t = sc["lifetime"]
    a = E * 2.0 / m[0]
    assert pos[0, 0] == pytest.approx(1e-4 + 50.0 * t + 0.5 * a * t * t, rel=1e-14)
    assert vel[0, 0] == pytest.approx(50.0 + a * t, rel=1e-14)


def test_dp5_fifth_order_convergence():
    # axial harmonic motion in the DC quadrupole: global error ~ h^5
    p, v, m, ch = _single([0, 0, 1e-3], [0, 0, 0])
    f = _fo(trap=1, Vrf=0.0, kappa=1.0, Vdc=5.0, z0=11e-3, r0=1.0)
    wz = math.sqrt(2 * E * 5.0 / 11e-3 ** 2 / m[0])
    T = 2 * 2 * math.pi / wz
    errs = []
    for nsteps in (200, 400):
        c = O.Cloud(p, v, m, ch, f, h0=T / nsteps, adaptive=False)
        for _ in range(nsteps):
            c.step()
        pos, _, sc = c.state()
        errs.append(abs(pos[0, 2] - 1e-3 * math.cos(wz * sc["lifetime"])))
    order = math.log2(errs[0] / errs[1])
    assert 4.7 < order < 5.5


def test_embedded_error_is_fourth_order():
    p, v, m, ch = _single([0, 0, 1e-3], [0, 0, 0])
    f = _fo(trap=1, Vrf=0.0, kappa=1.0, Vdc=5.0, z0=11e-3, r0=1.0)
    errs = []
    for h in (2e-7, 1e-7):
        c = O.Cloud(p, v, m, ch, f, h0=h, rtol=0.0, atol=1e-12, adaptive=False)
        errs.append(c.attempt())
    order = math.log2(errs[0] / errs[1])   # local error of the 4th-order pair: h^5
    assert 4.6 < order < 5.4


@pytest.mark.parametrize("err,reject,want_accept,want_scale", [
    (0.0, False, True, 10.0),
    (1.0, False, True, 0.9),
    (0.5 ** 5, False, True, 1.8),
    (1e-12, False, True, 10.0),      # clamp to maxscale
    (0.5 ** 5, True, True, 1.0),     # no growth right after a reject
    (32.0, False, False, 0.45),      # 0.9 * 32^-0.2
    (1e10, False, False, 0.2),       # clamp to minscale
])
def test_controller_closed_forms(err, reject, want_accept, want_scale):
    ok, h, hn, eo, rj = O.controller(err, 1e-8, 1e-8, 1e-4, reject)
    assert ok == want_accept
    if ok:
        assert hn == pytest.approx(1e-8 * want_scale, rel=1e-14)
        assert h == 1e-8
        assert eo == max(err, 1e-4)
        assert rj is False
    else:
        assert h == pytest.approx(1e-8 * want_scale, rel=1e-14)
        assert rj is True


@pytest.mark.parametrize("beta", [0.04, 0.05, 0.08])
def test_pi_controller_exponents(beta):
    """PI control (the reference's documented switch, bin/ode.cpp:427-432):
    log(hnext/h) = log 0.9 - alpha log err + beta log errold with
    alpha = 0.2 - 0.75 beta, inside the clamps.  Scaling errold by k moves
    hnext by exactly k^beta, and scaling err by k moves it by k^-alpha; k is
    chosen so that each factor is 2.  A sign slip, beta on the wrong error or
    an alpha that ignores beta fails; so does a reject that uses beta."""
    alpha = 0.2 - 0.75 * beta
    err, errold = 0.5, 0.05
    k_old = 2.0 ** (1.0 / beta)            # errold x k_old  -> hnext x 2
    k_err = 2.0 ** (-1.0 / alpha)          # err x k_err     -> hnext x 2
    _, _, h0, _, _ = O.controller(err, 1e-8, 1e-8, errold / k_old ** 0.5, False, beta=beta)
    _, _, h1, _, _ = O.controller(err, 1e-8, 1e-8, errold * k_old ** 0.5, False, beta=beta)
    assert h1 / h0 == pytest.approx(2.0, rel=1e-13)
    _, _, h2, eo, _ = O.controller(err * k_err, 1e-8, 1e-8, errold, False, beta=beta)
    _, _, h3, _, _ = O.controller(err, 1e-8, 1e-8, errold, False, beta=beta)
    assert h2 / h3 == pytest.approx(2.0, rel=1e-13)
    assert eo == err * k_err                 # errold bookkeeping is beta-free
    assert 0.2 < h3 / 1e-8 < 10.0 and 0.2 < h2 / 1e-8 < 10.0   # inside the clamps
    # beta = 0 is the reference's controller exactly
    assert O.controller(err, 1e-8, 1e-8, errold, False) == O.controller(err, 1e-8, 1e-8, 0.37, False, beta=0.0)
    # a reject does not look at errold
    r1 = O.controller(4.0, 1e-8, 1e-8, 1e-4, False, beta=beta)
    r2 = O.controller(4.0, 1e-8, 1e-8, 0.9, False, beta=beta)
    assert r1[:3] == r2[:3] and r1[4] == r2[4] and r1[0] is False
    assert (r1[3], r2[3]) == (1e-4, 0.9)     # a reject leaves errold alone
    assert r1[1] / 1e-8 == pytest.approx(max(0.9 * 4.0 ** -alpha, 0.2), rel=1e-14)


def test_pi_controller_in_a_run_changes_the_step_sequence():
    # the same cloud with beta 0 and 0.08: both reach t_end, with different step counts
    rng = np.random.Generator(np.random.PCG64(3))
    n = 24
    pos = rng.normal(0, 2e-4, (n, 3)); vel = rng.normal(0, 100.0, (n, 3))
    m = O.mass_kg(np.full(n, 133.0)); ch = np.ones(n, np.int32)
    f = _fo(trap=1, Vrf=0.0, kappa=1.0, Vdc=5.0, z0=11e-3, r0=1.0, coulomb=1)
    runs = {}
    for beta in (0.0, 0.08):
        c = O.Cloud(pos, vel, m, ch, f, h0=1e-9, rtol=1e-9, atol=1e-9, beta=beta)
        c.advance(2e-6)
        p, v, sc = c.state()
        runs[beta] = (p, sc)
    # (positions are not compared: each run overshoots t_end by its own last step)
    assert runs[0.0][1]["steps"] != runs[0.08][1]["steps"]
    assert min(runs[0.0][1]["lifetime"], runs[0.08][1]["lifetime"]) >= 2e-6


def test_stepsize_underflow_raises():
    # harmonic motion with an unreachable tolerance: h shrinks below |x| DBL_EPSILON
    p, v, m, ch = _single([1e-3, 0, 1e-3], [100.0, 0, 0])
    f = _fo(trap=1, Vrf=0.0, kappa=1.0, Vdc=5.0, z0=11e-3, r0=1.0)
    c = O.Cloud(p, v, m, ch, f, h0=1e-6, rtol=0.0, atol=1e-300, adaptive=True)
    with pytest.raises(RuntimeError):
        c.step()


def test_penning_single_ion_eigenmodes():
    # pure magnetron and pure cyclotron orbits: u = x+iy = r exp(-i w t)
    m133 = O.mass_kg([133.0])[0]
    wc = E * 7.0 / m133
    wz2 = E * 10.0 / (m133 * 11e-3 ** 2)
    root = math.sqrt(wc * wc / 4 - wz2 / 2)
    f = _fo(trap=0, Bz=7.0, V0=10.0, d0=11e-3)
    for w, r, T in ((wc / 2 - root, 1e-3, 2e-4), (wc / 2 + root, 1e-5, 2e-6)):
        p, v, m, ch = _single([r, 0, 0], [0, -w * r, 0])
        c = O.Cloud(p, v, m, ch, f, h0=1e-9, rtol=1e-11, atol=1e-14, time_mode=O.TIME_STAGE)
        c.advance(T)
        pos, _, sc = c.state()
        t = sc["lifetime"]
        assert pos[0, 0] == pytest.approx(r * math.cos(w * t), abs=1e-6 * r)
        assert pos[0, 1] == pytest.approx(-r * math.sin(w * t), abs=1e-6 * r)


def test_penning_energy_and_canonical_momentum_conserved(rng):
    n = 24
    m = O.mass_kg(np.full(n, 133.0)); ch = np.ones(n, dtype=np.int32)
    pos = rng.normal(0, 2e-4, (n, 3)); vel = rng.normal(0, 50.0, (n, 3))
    f = _fo(trap=0, Bz=7.0, V0=10.0, d0=11e-3, coulomb=1)
    c = O.Cloud(pos, vel, m, ch, f, h0=1e-9, rtol=1e-10, atol=1e-13, time_mode=O.TIME_STAGE)
    cpen = 10.0 / (2 * 11e-3 ** 2)

    def invariants(P, V):
        kin = 0.5 * (m[:, None] * V * V).sum()
        trap = (E * cpen * (P[:, 2] ** 2 - 0.5 * (P[:, 0] ** 2 + P[:, 1] ** 2))).sum()
        d = P[:, None, :] - P[None, :, :]
        r = np.sqrt((d * d).sum(-1)); np.fill_diagonal(r, np.inf)
        coul = 0.5 * (KE * E * E / r).sum()
        Lz = (m * (P[:, 0] * V[:, 1] - P[:, 1] * V[:, 0])).sum() + \
            (E * 7.0 * (P[:, 0] ** 2 + P[:, 1] ** 2) / 2).sum()
        return kin + trap + coul, Lz

    E0, L0 = invariants(pos, vel)
    c.advance(5e-6)
    P, V, _ = c.state()
    E1, L1 = invariants(P, V)
    assert abs(E1 - E0) < 1e-7 * abs(E0)
    assert abs(L1 - L0) < 1e-7 * abs(L0)


def test_friction_closed_form():
    p, v, m, ch = _single([0, 0, 0], [100.0, -20.0, 5.0])
    f = _fo(trap=0, delta=1e3)
    c = O.Cloud(p, v, m, ch, f, h0=1e-7, rtol=1e-12, atol=1e-15, time_mode=O.TIME_STAGE)
    c.advance(2e-3)
    pos, vel, sc = c.state()
    t = sc["lifetime"]
    v0 = np.array([100.0, -20.0, 5.0])
    assert np.allclose(vel[0], v0 * math.exp(-1e3 * t), rtol=1e-9)
    assert np.allclose(pos[0], v0 * (1 - math.exp(-1e3 * t)) / 1e3, rtol=1e-9)


def test_dipole_drive_closed_form_needs_stage_time():
    # free ion driven by Ed cos(w t): x = A (1 - cos w t)/w^2 from rest
    p, v, m, ch = _single([0, 0, 0], [0, 0, 0])
    w = 2 * math.pi * 1e5
    f = _fo(trap=0, Ed=10.0, fd=1e5)
    A = E * 10.0 / m[0]
    c = O.Cloud(p, v, m, ch, f, h0=1e-8, rtol=1e-12, atol=1e-16, time_mode=O.TIME_STAGE)
    c.advance(3.3e-5)
    pos, vel, sc = c.state()
    t = sc["lifetime"]
    assert pos[0, 0] == pytest.approx(A * (1 - math.cos(w * t)) / w ** 2, rel=1e-8)
    assert vel[0, 0] == pytest.approx(A * math.sin(w * t) / w, rel=1e-8, abs=1e-9 * A / w)
    # frozen time (the reference's force.cpp:36) is only first order in time
    c2 = O.Cloud(p, v, m, ch, f, h0=1e-8, adaptive=False, time_mode=O.TIME_FROZEN)
    c2.advance(3.3e-5)
    pos2, _, _ = c2.state()
    assert abs(pos2[0, 0] - pos[0, 0]) > 1e-4 * abs(pos[0, 0])


def test_trap_ramp_matches_manual_interpolation(rng):
    n = 8
    m = O.mass_kg(np.full(n, 133.0)); ch = np.ones(n, dtype=np.int32)
    pos = rng.normal(0, 3e-4, (n, 3)); vel = rng.normal(0, 30.0, (n, 3))
    base = dict(trap=1, Vrf=50.0, Vdc=5.0, kappa=0.001, r0=0.04, z0=0.04, freq_rf=5e6)
    fr = _fo(ramp=1, ramp_t0=0.0, ramp_t1=1e-6, newVrf=20.0, newVdc=9.0, **base)
    a = O.Cloud(pos, vel, m, ch, fr, h0=1e-8, adaptive=False)
    a.advance(2e-7)
    b = O.Cloud(pos, vel, m, ch, _fo(**base), h0=1e-8, adaptive=False)
    for _ in range(20):
        t = b.state()[2]["lifetime"]
        ramp = (t - 0.0) / (1e-6 - 0.0)
        b.set_fields(_fo(**dict(base, Vrf=50.0 * (1 - ramp) + 20.0 * ramp,
                                Vdc=5.0 * (1 - ramp) + 9.0 * ramp)))
        b.step()
    assert np.array_equal(a.state()[0], b.state()[0])
