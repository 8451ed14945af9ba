"""GPU parity of the HIP engine against the CPU oracle and the reference's
golden vectors.  Every call goes through the C ABI of libsimbuca_hip.so.

Tolerances:
* FP64-exact Coulomb, force() without time-dependent cosines, stage
  combinations: bit-identical (same operations in the same order, IEEE
  sqrt/div on both sides, no FMA contraction).
* Terms with cos(): relative 1e-14 (device and host libm may differ by an ulp).
* Trajectories: the device sums the error norm in a different association
  than the serial CPU loop, so h differs in the last bits: relative 1e-9 on
  states after 200 steps, exact accept/reject sequence.
* Mixed precision Coulomb: normwise ||a - a64|| / ||a64|| <= 1e-6.
"""
import numpy as np
import pytest

import oracle as O
import simbuca_amd as S
from simbuca_amd import synth
from conftest import golden_fields, load_golden

pytestmark = pytest.mark.gpu


def _both_fields(**kw):
    return S.default_fields(**kw), O.fields(**kw)


def _rand_cloud(rng, n, sigma=5e-4, species=(133.0, 85.0), charges=(1, 1, 2, -1)):
    amu = rng.choice(np.array(species), n)
    ch = rng.choice(np.array(charges, dtype=np.int32), n).astype(np.int32)
    pos = rng.normal(0, sigma, (n, 3))
    vel = rng.normal(0, 137.0, (n, 3))
    return amu, O.mass_kg(amu), ch, pos, vel


def _rows_error(a, ref):
    return np.linalg.norm(a - ref) / np.linalg.norm(ref)


# ---------------------------------------------------------------------------
# force() and the Coulomb sum
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("n", [1, 2, 255, 256, 257, 1000])
def test_coulomb_fp64_exact_is_bit_identical(rng, n):
    amu, m, ch, pos, vel = _rand_cloud(rng, n)
    sf, of = _both_fields(trap=0, coulomb=1, coulomb_scale=1.0)
    with S.Engine(precision="fp64_exact") as e:
        e.add_cloud(pos, vel, m, ch, sf)
        f = e.eval_coulomb(0, "fp64_exact")
    ref = O.coulomb_rows(pos, m, ch, 1.0, np.arange(n, dtype=np.int32))
    assert np.array_equal(f, ref)


@pytest.mark.parametrize("name", ["force_n500_trap_tof_coulomb.npz", "force_n257_coulomb_scaled.npz"])
def test_force_matches_reference_golden(name):
    g = load_golden(name)
    sf = golden_fields(g["fields"], S.default_fields)
    m = O.mass_kg(g["mass_amu"])
    with S.Engine(precision="fp64_exact") as e:
        e.add_cloud(g["pos"], g["vel"], m, g["charge"], sf, lifetime=float(g["lifetime"]))
        out = e.eval_force(0)
    ref = g["derivs"]
    if sf.trap and sf.Vrf != 0.0:
        # only cos(omega t) may differ by an ulp
        assert np.allclose(out, ref, rtol=1e-14, atol=1e-14 * np.abs(ref).max())
    else:
        assert np.array_equal(out, ref)


def test_force_bit_identical_without_cosines(rng):
    amu, m, ch, pos, vel = _rand_cloud(rng, 700)
    kw = dict(trap=1, Vrf=0.0, Vdc=4.0, kappa=0.3, r0=0.01, z0=0.02, tof=1, E_kick=-2.0,
              coulomb=1, coulomb_scale=0.8, V0=10.0, d0=11e-3, Bz=7.0, delta=250.0)
    sf, of = _both_fields(**kw)
    with S.Engine(precision="fp64_exact") as e:
        e.add_cloud(pos, vel, m, ch, sf, lifetime=2.5e-6)
        out = e.eval_force(0)
    ref = O.force(pos, vel, m, ch, of, 2.5e-6)
    assert np.array_equal(out, ref)


def test_force_time_dependent_terms(rng):
    amu, m, ch, pos, vel = _rand_cloud(rng, 300)
    kw = dict(trap=1, Vrf=40.0, Vdc=4.0, kappa=0.3, r0=0.01, z0=0.02, freq_rf=1.3e6, coulomb=1,
              Ed=12.0, fd=9.4e2, phid=0.3, Uq=0.2, rq=9e-3, fq=8.08e5, phiq=-0.7)
    sf, of = _both_fields(**kw)
    with S.Engine(precision="fp64_exact") as e:
        e.add_cloud(pos, vel, m, ch, sf, lifetime=3.7e-6)
        out = e.eval_force(0)
    ref = O.force(pos, vel, m, ch, of, 3.7e-6)
    assert np.allclose(out, ref, rtol=1e-13, atol=1e-13 * np.abs(ref).max())


def test_field_extras_bit_identical(rng):
    # magnetic bottle and buffer gas by mobility have no cosine: bit-identical
    amu, m, ch, pos, vel = _rand_cloud(rng, 600)
    kw = dict(trap=0, coulomb=1, V0=10.0, d0=11e-3, Bz=7.0, B2=2.5e4, K0=1.7e-4, p_mbar=3e-3,
              T_gas=295.0, delta=40.0)
    sf, of = _both_fields(**kw)
    with S.Engine(precision="fp64_exact") as e:
        e.add_cloud(pos, vel, m, ch, sf, lifetime=1.1e-6)
        out = e.eval_force(0)
    assert np.array_equal(out, O.force(pos, vel, m, ch, of, 1.1e-6))


def _field_grid(nr=60, nz=91, dr=1e-4, z0=-4.5e-3, dz=1e-4, a=3e4, b=-2e4):
    """A nonlinear axisymmetric test field (E_r, E_z) [nr, nz] and its grid."""
    r = np.arange(nr) * dr
    z = z0 + np.arange(nz) * dz
    R, Z = np.meshgrid(r, z, indexing="ij")
    er = a * R * (1.0 + 300.0 * Z) + 1e6 * R ** 2
    ez = b * Z + 5e5 * R * R - 4e6 * Z ** 3
    return er, ez, dr, z0, dz


def test_anharmonic_trap_and_field_map_bit_identical(rng):
    # C4/C6 terms and a field map (inside and outside the grid) have no
    # cosine: the force is bit-identical to the oracle's, and a short adaptive
    # run follows the oracle's trajectory
    amu, m, ch, pos, vel = _rand_cloud(rng, 600, sigma=2e-3)
    kw = dict(trap=0, coulomb=1, V0=10.0, d0=11e-3, Bz=7.0, C4=0.03, C6=-0.02)
    sf, of = _both_fields(**kw)
    er, ez, dr, z0, dz = _field_grid()
    S.set_emap(sf, er, ez, dr, z0, dz, scale=0.7)
    O.set_emap(of, er, ez, dr, z0, dz, scale=0.7)
    with S.Engine(precision="fp64_exact", rtol=1e-9, atol=1e-9) as e:
        e.add_cloud(pos, vel, m, ch, sf, lifetime=2e-7)
        out = e.eval_force(0)
        for _ in range(25):
            e.step()
        gp, gv = e.state(0)
        st = e.status(0)
    assert np.array_equal(out, O.force(pos, vel, m, ch, of, 2e-7))
    c = O.Cloud(pos, vel, m, ch, of, lifetime=2e-7, h0=1e-9, rtol=1e-9, atol=1e-9)
    for _ in range(25):
        c.step()
    op, ov, sc = c.state()
    assert st["attempts"] == sc["attempts"]
    assert np.allclose(gp, op, rtol=1e-10, atol=1e-10 * np.abs(op).max())
    assert np.allclose(gv, ov, rtol=1e-10, atol=1e-10 * np.abs(ov).max())


def test_field_maps_and_swift_share_the_sample_buffer(rng):
    # two clouds, one with a SWIFT waveform and a map, one with a larger map;
    # then the first cloud's map is replaced: every force equals the oracle's
    sizes = (200, 300)
    clouds, fields = [], []
    g1 = _field_grid(nr=20, nz=31, dr=3e-4, z0=-4.5e-3, dz=3e-4)
    g2 = _field_grid(nr=70, nz=101, dr=1e-4, z0=-5e-3, dz=1e-4, a=-2e4, b=5e4)
    for k, n in enumerate(sizes):
        amu, m, ch, pos, vel = _rand_cloud(rng, n, sigma=1e-3)
        sf, of = _both_fields(trap=0, coulomb=1, V0=10.0, d0=11e-3, C4=0.01 * (k + 1))
        g = g1 if k == 0 else g2
        S.set_emap(sf, *g); O.set_emap(of, *g)
        if k == 0:
            w = np.sin(np.linspace(0, 9, 40))
            S.set_swift(sf, w, 0.0, 1e-8, 5e-3, kind=2); O.set_swift(of, w, 0.0, 1e-8, 5e-3, kind=2)
        clouds.append((pos, vel, m, ch)); fields.append((sf, of))
    with S.Engine(precision="fp64_exact") as e:
        for (pos, vel, m, ch), (sf, _) in zip(clouds, fields):
            e.add_cloud(pos, vel, m, ch, sf, lifetime=1.3e-7)
        for k in range(2):
            pos, vel, m, ch = clouds[k]
            ref = O.force(pos, vel, m, ch, fields[k][1], 1.3e-7)
            assert np.allclose(e.eval_force(k), ref, rtol=1e-13, atol=1e-13 * np.abs(ref).max()), k
        sf, of = _both_fields(trap=0, coulomb=1, V0=10.0, d0=11e-3, C6=0.05)
        S.set_emap(sf, *g2); O.set_emap(of, *g2)
        e.set_fields(0, sf)
        for k, f_of in ((0, of), (1, fields[1][1])):
            pos, vel, m, ch = clouds[k]
            ref = O.force(pos, vel, m, ch, f_of, 1.3e-7)
            assert np.array_equal(e.eval_force(k), ref), k
    # a malformed map is refused
    bad = S.set_emap(S.default_fields(trap=0), *g1)
    bad.emap_dz = 0.0
    with S.Engine(precision="mixed") as e:
        with pytest.raises(S.SimbucaError, match="field map"):
            e.add_cloud(*clouds[0], bad)


@pytest.mark.parametrize("order", [1, 2])
def test_excitation_menu_matches_oracle(rng, order):
    # octupole, rotating wall, axial drive and swept dipole/quadrupole drives
    amu, m, ch, pos, vel = _rand_cloud(rng, 400)
    kw = dict(trap=0, coulomb=1, V0=10.0, d0=11e-3, Bz=7.0, Ed=12.0, fd=9.4e2, fd_rate=3e9,
              phid=0.3, Uq=0.2, rq=9e-3, fq=8.08e5, fq_rate=-2e10, phiq=-0.7, Uo=0.5, ro=7e-3,
              fo=1.6e6, phio=0.2, rw_order=order, Uw=0.3, rw=6e-3, fw=2.5e5, phiw=1.1, Eax=3.0,
              fax=1.9e5, phiax=-0.4)
    sf, of = _both_fields(**kw)
    with S.Engine(precision="fp64_exact") as e:
        e.add_cloud(pos, vel, m, ch, sf, lifetime=4.3e-6)
        out = e.eval_force(0)
    ref = O.force(pos, vel, m, ch, of, 4.3e-6)
    # only the device and host cos/sin may differ, by an ulp
    assert np.allclose(out, ref, rtol=1e-13, atol=1e-13 * np.abs(ref).max())


def test_excitation_menu_trajectory_vs_oracle(rng):
    # stage-time DP5 with every new term on, in the mixed engine against the
    # FP64 oracle over 30 fixed steps
    amu, m, ch, pos, vel = _rand_cloud(rng, 300, species=(133.0, 85.0), charges=(1,))
    nu_c, _, nu_m, _ = synth.penning_frequencies(133.0)
    kw = dict(trap=0, coulomb=1, V0=10.0, d0=11e-3, Bz=7.0, B2=1e4, Ed=10.0, fd=nu_m, fd_rate=1e9,
              Uq=0.1, rq=11e-3, fq=nu_c, Uo=0.05, ro=11e-3, fo=2 * nu_c, rw_order=2, Uw=0.02,
              rw=11e-3, fw=1e4, Eax=1.0, fax=3e5, K0=1.6e-4, p_mbar=1e-4, T_gas=300.0)
    sf, of = _both_fields(**kw)
    with S.Engine(precision="mixed", adaptive=False, h0=2e-9, time_mode=S.TIME_STAGE) as e:
        e.add_cloud(pos, vel, m, ch, sf)
        e.advance(5.9e-8)
        gp, gv = e.state(0)
        st = e.status(0)
    c = O.Cloud(pos, vel, m, ch, of, h0=2e-9, adaptive=False, time_mode=O.TIME_STAGE)
    for _ in range(st["steps"]):
        c.step()
    op, ov, sc = c.state()
    assert st["steps"] == 30 and st["lifetime"] == sc["lifetime"]
    assert np.allclose(gp, op, rtol=0, atol=1e-12 * np.abs(op).max())
    assert np.allclose(gv, ov, rtol=0, atol=1e-8 * np.abs(ov).max())


@pytest.mark.parametrize("kind", [1, 2])
def test_swift_bit_identical(rng, kind):
    # linear interpolation has no transcendental: bit-identical to the oracle,
    # inside and outside the sampled window
    amu, m, ch, pos, vel = _rand_cloud(rng, 500)
    samples = rng.normal(0, 0.3, 257)
    sf = S.set_swift(S.default_fields(trap=0, coulomb=1, Bz=7.0, V0=10.0, d0=11e-3),
                     samples, 2e-7, 3e-9, 7e-3, kind)
    of = O.set_swift(O.fields(trap=0, coulomb=1, Bz=7.0, V0=10.0, d0=11e-3),
                     samples, 2e-7, 3e-9, 7e-3, kind)
    for t in (1e-7, 2e-7, 4.123e-7, 9.68e-7, 2e-7 + 256 * 3e-9, 1.2e-6):
        with S.Engine(precision="fp64_exact") as e:
            e.add_cloud(pos, vel, m, ch, sf, lifetime=t)
            out = e.eval_force(0)
        assert np.array_equal(out, O.force(pos, vel, m, ch, of, t)), t


def test_swift_set_fields_moves_samples(rng):
    # two clouds; replacing cloud 0's waveform with a longer one reallocates the
    # device buffer and moves cloud 1's samples: both must follow the oracle
    clouds = [_rand_cloud(rng, n, species=(133.0,), charges=(1,)) for n in (120, 90)]
    w0, w1, w0b = rng.normal(0, 1, 50), rng.normal(0, 1, 80), rng.normal(0, 1, 400)
    base = dict(trap=0, coulomb=1, Bz=7.0, V0=10.0, d0=11e-3)
    sfs = [S.set_swift(S.default_fields(**base), w0, 0.0, 2e-9, 5e-3, 1),
           S.set_swift(S.default_fields(**base), w1, 0.0, 2e-9, 5e-3, 2)]
    ofs = [O.set_swift(O.fields(**base), w0b, 0.0, 2e-9, 5e-3, 1),
           O.set_swift(O.fields(**base), w1, 0.0, 2e-9, 5e-3, 2)]
    with S.Engine(precision="fp64_exact", adaptive=False, h0=1e-9, time_mode=S.TIME_STAGE) as e:
        for (_, m, ch, pos, vel), sf in zip(clouds, sfs):
            e.add_cloud(pos, vel, m, ch, sf)
        e.step()
        e.set_fields(0, S.set_swift(S.default_fields(**base), w0b, 0.0, 2e-9, 5e-3, 1))
        for _ in range(5):
            e.step()
        got = [e.state(i) for i in range(2)]
    for ci, ((_, m, ch, pos, vel), of) in enumerate(zip(clouds, ofs)):
        first = O.set_swift(O.fields(**base), w0, 0.0, 2e-9, 5e-3, 1) if ci == 0 else of
        c = O.Cloud(pos, vel, m, ch, first, h0=1e-9, adaptive=False, time_mode=O.TIME_STAGE)
        c.step()
        c.set_fields(of)
        for _ in range(5):
            c.step()
        op, ov, _ = c.state()
        assert np.array_equal(got[ci][0], op) and np.array_equal(got[ci][1], ov), ci


def test_field_parameters_are_validated(rng):
    amu, m, ch, pos, vel = _rand_cloud(rng, 10)
    with S.Engine(precision="fp64_exact") as e:
        with pytest.raises(S.SimbucaError, match="rw_order"):
            e.add_cloud(pos, vel, m, ch, S.default_fields(trap=0, Uw=1.0, rw_order=3))
        with pytest.raises(S.SimbucaError, match="radii"):
            e.add_cloud(pos, vel, m, ch, S.default_fields(trap=0, Uo=1.0, ro=0.0))
        with pytest.raises(S.SimbucaError, match="T_gas"):
            e.add_cloud(pos, vel, m, ch, S.default_fields(trap=0, K0=1e-4, p_mbar=1.0, T_gas=0.0))
        with pytest.raises(S.SimbucaError, match="SWIFT"):
            e.add_cloud(pos, vel, m, ch, S.set_swift(S.default_fields(trap=0), [1.0], 0.0, 1e-9, 1.0))
        with pytest.raises(S.SimbucaError, match="SWIFT"):
            e.add_cloud(pos, vel, m, ch, S.set_swift(S.default_fields(trap=0), [1.0, 2.0], 0.0, 0.0, 1.0))
        e.add_cloud(pos, vel, m, ch, S.default_fields(trap=0, Uw=1.0, rw_order=1))


@pytest.fixture(scope="module")
def exact_65536():
    """FP64-exact Coulomb sum of the 65,536-ion cloud, summed in the caller's
    order (FP64-exact engines never reorder), checked against oracle rows."""
    c = synth.workload(3).clouds[0]
    rows = np.random.Generator(np.random.PCG64(7)).choice(c.n, 192, replace=False).astype(np.int32)
    ref = O.coulomb_rows(c.pos, c.mass, c.charge, 1.0, rows)
    with S.Engine(precision="fp64_exact") as e:
        e.add_cloud(c.pos, c.vel, c.mass, c.charge, c.fields)
        exact = e.eval_coulomb(0, "fp64_exact")
    assert np.array_equal(exact[rows], ref)
    return exact


@pytest.mark.parametrize("precision,local,tol", [("fp64", 0, 1e-13), ("fp64_rsq", 0, 1e-14),
                                                 ("mixed", 0, 1e-6),
                                                 ("mixed", 1, 1e-6), ("mixed_kahan", 0, 1e-6),
                                                 ("mixed_kahan", 1, 1e-6)])
def test_coulomb_precisions_at_65536(exact_65536, precision, local, tol):
    c = synth.workload(3).clouds[0]
    with S.Engine(precision=precision) as e:
        e.add_cloud(c.pos, c.vel, c.mass, c.charge, c.fields)
        e.set_launch(4, 4, 0, local)
        f = e.eval_coulomb(0, precision)
    err = _rows_error(f, exact_65536)
    assert err <= tol, err
    # Newton's third law holds to the same accuracy (equal masses and charges)
    assert np.linalg.norm(f.sum(0)) <= 10 * tol * np.abs(f).sum(0).max()


@pytest.mark.parametrize("index", [2, 4, 5])
def test_coulomb_at_full_workload_sizes(index):
    """The other benchmark set-ups at full size: 10,000 ions of two species,
    the 256-cloud ensemble, 262,144 ions of three species (Kahan).  The
    workload's precision, with its automatic launch shape, against the FP64
    sum, which is pinned to the oracle on sampled rows of the largest cloud;
    Newton's third law holds in every cloud."""
    w = synth.workload(index)
    print(f"full-size Coulomb, workload {index}: {len(w.clouds)} clouds, "
          f"{sum(c.n for c in w.clouds)} ions", flush=True)
    with S.Engine(precision=w.precision) as e:
        for c in w.clouds:
            e.add_cloud(c.pos, c.vel, c.mass, c.charge, c.fields)
        mixed = [e.eval_coulomb(k, w.precision) for k in range(len(w.clouds))]
        fp64 = [e.eval_coulomb(k, "fp64") for k in range(len(w.clouds))]
    big = int(np.argmax([c.n for c in w.clouds]))
    c = w.clouds[big]
    rows = np.random.Generator(np.random.PCG64(index)).choice(c.n, 64, replace=False).astype(np.int32)
    ref = O.coulomb_rows(c.pos, c.mass, c.charge, c.fields.coulomb_scale, rows)
    assert _rows_error(fp64[big][rows], ref) < 1e-13
    errs = [_rows_error(a, b) for a, b in zip(mixed, fp64)]
    assert max(errs) <= 1e-6, max(errs)
    assert _rows_error(np.concatenate(mixed), np.concatenate(fp64)) <= 1e-6
    for cl, a in zip(w.clouds, mixed):
        if cl.n < 2:
            continue
        ma = a * cl.mass[:, None]
        assert np.linalg.norm(ma.sum(0)) <= 1e-5 * np.abs(ma).sum(0).max()


def test_local_coords_keep_the_accuracy(exact_65536):
    # far sub-tiles from FP32 offsets: same normwise error as FP64-formed
    # separations everywhere (measured 5.6e-8 both), well inside 2x
    c = synth.workload(3).clouds[0]
    errs = {}
    with S.Engine(precision="mixed") as e:
        e.add_cloud(c.pos, c.vel, c.mass, c.charge, c.fields)
        for local in (0, 1):
            e.set_launch(4, 4, 0, local)
            errs[local] = _rows_error(e.eval_coulomb(0, "mixed"), exact_65536)
    assert errs[1] <= 2.0 * errs[0], errs


@pytest.mark.parametrize("precision", ["mixed", "mixed_kahan"])
@pytest.mark.parametrize("unit", [True, False])
@pytest.mark.parametrize("tpt,groups", [(1, 4), (2, 4), (4, 4), (2, 1), (8, 4), (8, 1)])
def test_local_coords_close_pairs_take_the_fp64_separation(rng, precision, unit, tpt, groups):
    # 16 pairs of ions 20 nm apart in a 5,003-ion cloud: each partner's pull
    # outweighs the rest of the cloud about 1e5 times, so the per-ion error of
    # these 32 ions is the error of one near pair.  Formed from FP32 offsets
    # to a block origin some 1e-4 m away, that separation would err by about
    # 1e-4 of itself; formed in FP64 it errs by about 1e-7, and the mixed
    # kernel's softening (1e-23 m^2 added to r^2) moves it by 4e-8.  A source
    # put on the wrong path, or a lost or doubled staging slot, fails this test.
    n = 5003
    pos = rng.normal(0, 5e-4, (n, 3))
    pos[4000:4016] = pos[100:116] + np.array([2e-8, 0.0, 0.0])
    close = np.r_[100:116, 4000:4016]
    amu = np.full(n, 133.0)
    ch = np.ones(n, np.int32) if unit else rng.choice(np.array([1, 2, 3], np.int32), n)
    sf, _ = _both_fields(trap=0, coulomb=1)
    with S.Engine(precision=precision) as e:
        e.add_cloud(pos, np.zeros((n, 3)), O.mass_kg(amu), ch, sf)
        exact = e.eval_coulomb(0, "fp64_exact")
        for local in (1, 0):
            e.set_launch(tpt, groups, 0, local)
            a = e.eval_coulomb(0, precision)
            assert _rows_error(a, exact) < 1e-6, local
            per_ion = np.linalg.norm(a[close] - exact[close], axis=1) / np.linalg.norm(exact[close], axis=1)
            assert per_ion.max() < 1e-6, (local, per_ion.max())


def _sizes(kind):
    if kind == "small ensemble":      # 512 clouds of 64-512 ions: 4e7 pairs
        r = np.random.default_rng(11)
        return np.round(np.exp(r.uniform(np.log(64), np.log(512), 512))).astype(int).tolist()
    if kind == "ensemble":            # the 256-cloud set-up, 64-4,096 ions
        return [c.n for c in synth.workload(4).clouds]
    return [int(kind)]


@pytest.mark.parametrize("kind,precision,want", [
    ("500", "mixed", (1, 4, 0, 1)),
    ("6000", "mixed", (1, 4, 0, 1)),
    ("10000", "mixed", (2, 4, 1, 1)),
    ("20000", "mixed", (2, 4, 1, 1)),
    ("40000", "mixed", (4, 4, 1, 1)),
    ("small ensemble", "mixed", (1, 4, 0, 1)),
    ("ensemble", "mixed", (2, 4, 1, 1)),
    ("10000", "fp64", (1, 4, 0, 0)),
    ("40000", "fp64", (4, 4, 0, 0)),
    ("ensemble", "fp64", (2, 4, 0, 0)),
])
def test_automatic_launch_shape(kind, precision, want):
    """The automatic launch shape as sbk_set_launch documents it: 4 targets
    per thread from a pair-weighted mean cloud size of 32,768 ions, else 2
    from 2.5e8 pairs or with local coordinates, else 1; local coordinates
    (mixed, spatial order) from 8e7 pairs and a pair-weighted mean size of
    2,048 ions; the streamed kernel for the mixed precisions with 4 source
    groups.  The resolved slice count is that of the work list."""
    sizes = _sizes(kind)
    print(f"automatic launch: {kind} ({len(sizes)} clouds, {sum(sizes)} ions), {precision}", flush=True)
    f = synth.penning_fields()
    with S.Engine(precision=precision) as e:
        for i, n in enumerate(sizes):
            c = synth.gaussian_cloud(n, 1e-3, 300.0, seed=100 + i)
            e.add_cloud(c.pos, c.vel, c.mass, c.charge, f)
        tpt, groups, slices, local, kernel = e.launch(precision)
        assert (tpt, groups, local, kernel) == want
        assert slices >= 1
        # an explicit shape is reported back as set (groups 1: the tiled kernel)
        e.set_launch(1, 1, 3, 0)
        t1, g1, s1, l1, k1 = e.launch(precision)
        assert (t1, g1, l1, k1) == (1, 1, 0, 0)
        assert s1 >= 3 if min(sizes) >= 3 * 1024 else s1 >= 1


def test_automatic_launch_shape_headline():
    """65,536 ions, mixed: 4 targets per thread, 4 source groups, 32 slices
    (eight rounds of the resident work-groups), local coordinates, the
    streamed kernel: the shape the shipped table also records."""
    c = synth.workload(3).clouds[0]
    with S.Engine(precision="mixed") as e:
        e.add_cloud(c.pos, c.vel, c.mass, c.charge, c.fields)
        assert e.launch("mixed") == (4, 4, 32, 1, 1)
        assert e.launch("fp64_exact")[:3] == (1, 1, 1)


@pytest.mark.parametrize("tpt", [1, 2, 4, 8])
@pytest.mark.parametrize("groups", [1, 4])
@pytest.mark.parametrize("slices", [0, 1, 3, 7])
@pytest.mark.parametrize("local", [0, 1])
def test_launch_shapes_agree(rng, tpt, groups, slices, local):
    amu, m, ch, pos, vel = _rand_cloud(rng, 3001, sigma=1e-3)
    sf, _ = _both_fields(trap=0, coulomb=1)
    with S.Engine(precision="mixed") as e:
        e.add_cloud(pos, vel, m, ch, sf)
        e.set_launch(tpt, groups, slices, local)
        mixed = e.eval_coulomb(0, "mixed")
        fp64 = e.eval_coulomb(0, "fp64")
        exact = e.eval_coulomb(0, "fp64_exact")
    assert _rows_error(fp64, exact) < 1e-13
    assert _rows_error(mixed, exact) < 1e-6
    with S.Engine(precision="fp64_rsq") as e:
        e.add_cloud(pos, vel, m, ch, sf)
        e.set_launch(tpt, groups, slices, 0)
        rsq = e.eval_coulomb(0, "fp64_rsq")
    assert _rows_error(rsq, exact) < 1e-14


@pytest.mark.parametrize("n", [1, 2, 257, 4099])
def test_fp64_rsq_against_the_oracle(rng, n):
    """The FP64 kernel with the hardware |r|^-3 estimate and its second-order
    correction: within 1e-14 normwise of the oracle's FP64 sum
    (bin/force.cpp:147-175), charges and masses mixed; the self pair adds 0."""
    amu, m, ch, pos, vel = _rand_cloud(rng, n, sigma=1e-3)
    sf, of = _both_fields(trap=0, coulomb=1)
    ref = O.force(pos, vel, m, ch, of, 0.0)
    with S.Engine(precision="fp64_rsq") as e:
        e.add_cloud(pos, vel, m, ch, sf)
        f = e.eval_force(0)
        st = e.status(0)
    assert np.isfinite(f).all()
    if n == 1:
        assert np.array_equal(f, ref)
        return
    assert _rows_error(f, ref) < 1e-14, _rows_error(f, ref)
    assert st["error"] == 0


@pytest.mark.parametrize("precision", ["mixed", "mixed_kahan"])
@pytest.mark.parametrize("unit", [True, False])
def test_stream_kernel_equals_tiled_kernel(rng, precision, unit):
    """The streamed kernel gives wave w the same sub-tiles the tiled kernel's
    wave group w takes (w, w+4, ... of each slice), stages them in the same
    order, folds them in the same order and combines the four waves in the
    same order: its sums are the tiled kernel's bit for bit, for every shape
    they share (the streamed kernel's automatic split with 4 targets per
    thread may choose half-size target blocks, so there both are checked
    against the FP64-exact sum), and within 1e-6 of the FP64-exact sum.  Clouds of uneven
    sizes (ragged last tiles, a 3-ion cloud) share the launch; with 4 targets
    per thread their last blocks fill 3 (3001, 700), 1 (64, 3) and 2 (356)
    sets of 64 targets, which the streamed kernel runs with the instantiation
    for those sets."""
    sizes = [3001, 64, 3, 700, 356]
    sf, _ = _both_fields(trap=0, coulomb=1, V0=10.0, d0=11e-3, Bz=7.0)
    clouds = []
    for n in sizes:
        amu, m, ch, pos, vel = _rand_cloud(rng, n, sigma=1e-3)
        if unit:
            ch = np.ones(n, np.int32)
        clouds.append((pos, vel, m, ch))

    def engine():
        e = S.Engine(precision=precision, time_mode=S.TIME_STAGE, rtol=1e-9, atol=1e-9)
        for pos, vel, m, ch in clouds:
            e.add_cloud(pos, vel, m, ch, sf)
        return e

    with engine() as e:
        exact = [e.eval_coulomb(c, "fp64_exact") for c in range(len(sizes))]
        for tpt in (1, 2, 4):
            for slices in (0, 1, 3):
                for local in (0, 1):
                    print(f"stream vs tiled {precision} unit={unit} tpt={tpt} slices={slices} "
                          f"local={local}", flush=True)
                    e.set_launch(tpt, 4, slices, local)
                    e.set_coulomb_kernel("tiled")
                    tiled = [e.eval_coulomb(c, precision) for c in range(len(sizes))]
                    e.set_coulomb_kernel("stream")
                    stream = [e.eval_coulomb(c, precision) for c in range(len(sizes))]
                    for c in range(len(sizes)):
                        assert _rows_error(stream[c], exact[c]) < 1e-6
                        if tpt == 4 and slices == 0:
                            # the streamed kernel's automatic list gives clouds whose items
                            # would be large half-size target blocks (other boxes and slices)
                            assert _rows_error(tiled[c], exact[c]) < 1e-6
                        else:
                            assert np.array_equal(stream[c], tiled[c]), (tpt, slices, local, c)
    # whole adaptive attempts replayed from the graph: the same states bit for bit
    states = {}
    for kern in ("tiled", "stream"):
        with engine() as e:
            e.set_launch(2, 4, 0, 1)
            e.set_coulomb_kernel(kern)
            e.attempts(3)
            e.synchronize()
            states[kern] = [e.state(c) for c in range(len(sizes))]
    for (pt, vt), (ps, vs) in zip(states["tiled"], states["stream"]):
        assert np.array_equal(pt, ps) and np.array_equal(vt, vs)


def test_autotune_table_round_trip(rng, tmp_path):
    amu, m, ch, pos, vel = _rand_cloud(rng, 2000, sigma=1e-3)
    sf, _ = _both_fields(trap=0, coulomb=1)
    table = str(tmp_path / "tune.tsv")
    with S.Engine(precision="mixed") as e:
        e.add_cloud(pos, vel, m, ch, sf)
        shape = e.autotune("mixed", table)
        ref = e.eval_coulomb(0, "mixed")
    lines = open(table).read().strip().splitlines()
    # tiled: local coords x R x Q x slice rounds; streamed: local coords x R (Q = 4) x rounds
    assert len(lines) == 2 * 3 * 2 * 6 + 2 * 3 * 6
    assert {line.split()[7] for line in lines} == {"0", "1"}
    assert {line.split()[8] for line in lines} == {"0", "1"}   # both kernels
    assert len(shape) == 5 and shape[4] in (0, 1)
    with S.Engine(precision="mixed") as e:
        e.add_cloud(pos, vel, m, ch, sf)
        assert e.load_tuning(table, "mixed")
        assert not e.load_tuning(table, "fp64")        # no FP64 lines in the table
        again = e.eval_coulomb(0, "mixed")
    assert np.array_equal(again, ref)                  # same shape -> same sums


def test_spatial_order_is_invisible_to_the_caller(rng):
    # mixed engines keep each cloud in a spatial order on the device; states
    # and forces come back in the caller's order, and re-sorting an unmoved
    # cloud changes nothing at all
    amu, m, ch, pos, vel = _rand_cloud(rng, 5000, sigma=1e-3)
    kw = dict(trap=0, coulomb=1, V0=10.0, d0=11e-3, Bz=7.0)
    sf, of = _both_fields(**kw)
    assert not np.array_equal(S.spatial_order(pos), np.arange(5000))
    with S.Engine(precision="mixed") as e:
        e.add_cloud(pos, vel, m, ch, sf)
        e.set_launch(4, 4, 0, 1)
        gp, gv = e.state(0)
        f = e.eval_force(0)
        c = e.eval_coulomb(0, "mixed")
        e.reorder()
        gp2, gv2 = e.state(0)
        f2 = e.eval_force(0)
    assert np.array_equal(gp, pos) and np.array_equal(gv, vel)
    assert np.array_equal(gp2, pos) and np.array_equal(gv2, vel)
    assert np.array_equal(f, f2)
    ref = O.coulomb_rows(pos, m, ch, 1.0, np.arange(5000, dtype=np.int32))
    assert _rows_error(c, ref) <= 1e-6
    assert _rows_error(f, O.force(pos, vel, m, ch, of, 0.0)) <= 1e-6


@pytest.mark.parametrize("precision", ["mixed", "mixed_kahan"])
def test_reorder_during_a_run(rng, precision):
    # re-sorting after every batch of attempts (ions move) against never
    # re-sorting, over 40 fixed 1 ns steps (fixed steps: an adaptive run ends
    # where its last step lands, which rounding moves): same Coulomb
    # evaluations (the stage-1 derivative moves with its ion), and states
    # equal up to the mixed precision's summation order, which moves a
    # position by ~|a| * 5e-8 * t^2 / 2 ~ 1e-15 m here
    amu, m, ch, pos, vel = _rand_cloud(rng, 3000, sigma=3e-4, species=(133.0,), charges=(1,))
    kw = dict(trap=0, coulomb=1, V0=10.0, d0=11e-3, Bz=7.0, Ed=20.0, fd=3e5)
    sf, of = _both_fields(**kw)
    res = []
    for interval in (0, 1):
        with S.Engine(precision=precision, adaptive=False, h0=1e-9, time_mode=S.TIME_STAGE,
                      reorder_interval=interval) as e:
            e.add_cloud(pos, vel, m, ch, sf)
            e.set_launch(1, 4, 0, 1)
            e.advance(3.95e-8)
            res.append((*e.state(0), e.status(0)))
    (p0, v0, s0), (p1, v1, s1) = res
    assert s0["attempts"] == s1["attempts"] == 40
    assert s0["coulomb_evals"] == s1["coulomb_evals"] == 7 + 6 * 39
    assert s0["lifetime"] == s1["lifetime"]
    assert np.allclose(p1, p0, rtol=0, atol=1e-13)
    assert np.allclose(v1, v0, rtol=0, atol=1e-5)
    # and both follow the FP64 oracle to the mixed precision's accuracy
    c = O.Cloud(pos, vel, m, ch, of, h0=1e-9, adaptive=False, time_mode=O.TIME_STAGE)
    for _ in range(40):
        c.step()
    op, ov, _ = c.state()
    assert np.allclose(p1, op, rtol=0, atol=1e-13)
    assert np.allclose(v1, ov, rtol=0, atol=1e-5)


def test_reorder_interval_counts_launches_not_cloud_attempts(rng):
    """reorder_interval counts attempt launches (one attempt of every active
    cloud), not attempts summed over the clouds: 8 clouds, fixed steps, 64
    attempts each and an interval of 16 allow at most 64 / 16 = 4 re-sorts
    (a sum over the clouds would re-sort at every check of the run)."""
    sf, _ = _both_fields(trap=0, coulomb=1, V0=10.0, d0=11e-3, Bz=7.0)
    with S.Engine(precision="mixed", adaptive=False, h0=1e-9, reorder_interval=16) as e:
        for _ in range(8):
            amu, m, ch, pos, vel = _rand_cloud(rng, 300)
            e.add_cloud(pos, vel, m, ch, sf)
        e.advance(63.5e-9)
        e.synchronize()
        assert all(e.status(i)["attempts"] == 64 for i in range(8))
        assert 1 <= e.stats()["reorders"] <= 4


def test_ensemble_advance_equals_clouds_alone(rng):
    """An ensemble advanced together against each cloud advanced alone, FP64
    exact (whose sums do not depend on the work list), bit for bit.  The clouds
    need different numbers of attempts (the dense 900-ion one the most), so
    during the run the work list is planned again for the clouds still
    active; a cloud dropped from it, or a stale partial sum, would show."""
    # (ions, sigma, Coulomb term): a single ion and a cloud without Coulomb
    # term are in the work list's mask as any other cloud
    specs = [(900, 1e-4, 1), (40, 5e-4, 1), (200, 5e-4, 1), (500, 3e-4, 1), (1, 5e-4, 1),
             (300, 3e-4, 0)]
    clouds = [_rand_cloud(rng, n, sigma=sg) for n, sg, _ in specs]
    fields = [_both_fields(trap=0, coulomb=cb, V0=10.0, d0=11e-3, Bz=7.0)[0] for _, _, cb in specs]
    kw = dict(precision="fp64_exact", rtol=1e-9, atol=1e-9, h0=1e-9, time_mode=S.TIME_STAGE)
    with S.Engine(**kw) as e:
        for (_, m, ch, pos, vel), sf in zip(clouds, fields):
            e.add_cloud(pos, vel, m, ch, sf)
        e.advance(2e-7)
        together = [(*e.state(i), e.status(i)["steps"], e.status(i)["attempts"]) for i in range(len(specs))]
    att = [t[3] for t in together]
    assert att[0] == max(att) and att[0] > 1.5 * min(att), att
    for (_, m, ch, pos, vel), sf, (p0, v0, steps, attempts) in zip(clouds, fields, together):
        with S.Engine(**kw) as e:
            e.add_cloud(pos, vel, m, ch, sf)
            e.advance(2e-7)
            p1, v1 = e.state(0)
            assert (e.status(0)["steps"], e.status(0)["attempts"]) == (steps, attempts)
        assert np.array_equal(p0, p1) and np.array_equal(v0, v1)


def test_reorder_of_an_ensemble_keeps_every_cloud(rng):
    """A re-sort of an ensemble large enough that its clouds are bisected on
    several host threads (24 clouds, about 20,000 ions), after the ions have
    moved: every cloud's state, returned in the caller's order, is unchanged
    bit for bit, and its forces (charges and masses follow their ions) agree
    with those before the re-sort to the mixed precision's rounding."""
    sizes = [int(x) for x in rng.integers(300, 1400, 24)]
    assert sum(sizes) >= 16384
    clouds = [_rand_cloud(rng, n, sigma=4e-4, species=(133.0, 85.0), charges=(1, 2)) for n in sizes]
    sf, _ = _both_fields(trap=0, coulomb=1, V0=10.0, d0=11e-3, Bz=7.0)
    with S.Engine(precision="mixed", reorder_interval=0) as e:
        for _, m, ch, pos, vel in clouds:
            e.add_cloud(pos, vel, m, ch, sf)
        e.advance(1e-6)             # the ions move about a third of the cloud's width
        before = [(*e.state(i), e.eval_force(i), e.eval_coulomb(i, "mixed"))
                  for i in range(len(sizes))]
        e.reorder()
        assert e.stats()["reorders"] >= 1
        changed = 0
        for i, (p0, v0, f0, c0) in enumerate(before):
            p1, v1 = e.state(i)
            assert np.array_equal(p0, p1) and np.array_equal(v0, v1), i
            assert _rows_error(e.eval_force(i), f0) < 1e-6, i
            changed += not np.array_equal(e.eval_coulomb(i, "mixed"), c0)
        assert changed > 0          # the new order sums in another order


def test_multi_cloud_equals_single_clouds(rng):
    sizes = [64, 1, 300, 0, 1029]
    clouds = [_rand_cloud(rng, n) for n in sizes]
    sf, _ = _both_fields(trap=0, coulomb=1, V0=10.0, d0=11e-3, Bz=7.0)
    with S.Engine(precision="fp64_exact") as e:
        for _, m, ch, pos, vel in clouds:
            e.add_cloud(pos, vel, m, ch, sf)
        outs = [e.eval_force(i) for i in range(len(sizes))]
    for (amu, m, ch, pos, vel), out in zip(clouds, outs):
        if len(m) == 0:
            assert out.shape == (0, 3)
            continue
        ref = O.force(pos, vel, m, ch, O.fields(trap=0, coulomb=1, V0=10.0, d0=11e-3, Bz=7.0), 0.0)
        assert np.array_equal(out, ref)


# ---------------------------------------------------------------------------
# Dormand-Prince steps
# ---------------------------------------------------------------------------

def _golden_engine(g, use_graph):
    sf = golden_fields(g["fields"], S.default_fields)
    m = O.mass_kg(g["mass_amu"])
    e = S.Engine(precision="fp64_exact", adaptive=bool(g["adaptive"]), use_graph=use_graph,
                 rtol=float(g["rtol"]), atol=float(g["atol"]), h0=float(g["h0"]),
                 time_mode=S.TIME_FROZEN)
    e.add_cloud(g["pos"], g["vel"], m, g["charge"], sf, lifetime=float(g["lifetime"]))
    return e


def test_fixed_steps_match_reference_golden():
    g = load_golden("steps_fixed_rf_n64.npz")
    with _golden_engine(g, use_graph=False) as e:
        for _ in range(int(g["nsteps"])):
            e.step()
        pos, vel = e.state(0)
        st = e.status(0)
    assert st["lifetime"] == g["scalars"][0]
    # only the RF cosine may differ by an ulp
    assert np.allclose(pos, g["pos_out"], rtol=1e-13, atol=1e-13 * np.abs(g["pos_out"]).max())
    assert np.allclose(vel, g["vel_out"], rtol=1e-12, atol=1e-12 * np.abs(g["vel_out"]).max())


@pytest.mark.parametrize("use_graph", [False, True])
def test_adaptive_steps_match_reference_golden(use_graph):
    g = load_golden("steps_adaptive_n64.npz")
    with _golden_engine(g, use_graph=use_graph) as e:
        hist = []
        for _ in range(int(g["nsteps"])):
            hist.append(e.status(0)["h"])
            e.step()
        pos, vel = e.state(0)
        st = e.status(0)
    assert st["attempts"] == g["scalars"][5]          # same accept/reject sequence
    assert np.allclose(np.array(hist), g["hist"][:, 0], rtol=1e-9, atol=0)
    assert st["lifetime"] == pytest.approx(g["scalars"][0], rel=1e-9)
    assert np.allclose(pos, g["pos_out"], rtol=1e-8, atol=1e-8 * np.abs(g["pos_out"]).max())
    assert np.allclose(vel, g["vel_out"], rtol=1e-8, atol=1e-8 * np.abs(g["vel_out"]).max())


@pytest.mark.parametrize("beta", [0.04, 0.08])
def test_pi_controller_matches_oracle(beta):
    """PI control (sbk_config.beta, the reference's documented switch,
    bin/ode.cpp:427-432) on the golden 64-ion adaptive cloud: 200 steps with
    the same accept/reject sequence and h history as the oracle, whose
    controller takes the same beta."""
    g = load_golden("steps_adaptive_n64.npz")
    sf = golden_fields(g["fields"], S.default_fields)
    of = golden_fields(g["fields"], O.fields)
    m = O.mass_kg(g["mass_amu"])
    kw = dict(rtol=float(g["rtol"]), atol=float(g["atol"]), h0=float(g["h0"]))
    c = O.Cloud(g["pos"], g["vel"], m, g["charge"], of, lifetime=float(g["lifetime"]),
                adaptive=True, beta=beta, **kw)
    with S.Engine(precision="fp64_exact", adaptive=True, time_mode=S.TIME_FROZEN, beta=beta,
                  **kw) as e:
        e.add_cloud(g["pos"], g["vel"], m, g["charge"], sf, lifetime=float(g["lifetime"]))
        gh, oh = [], []
        for _ in range(200):
            gh.append(e.status(0)["h"])
            oh.append(c.state()[2]["h"])
            e.step()
            c.step()
        pos, vel = e.state(0)
        st = e.status(0)
    op, ov, sc = c.state()
    assert st["attempts"] == sc["attempts"] and st["steps"] == sc["steps"] == 200
    assert np.allclose(gh, oh, rtol=1e-9, atol=0)
    assert st["lifetime"] == pytest.approx(sc["lifetime"], rel=1e-9)
    assert np.allclose(pos, op, rtol=1e-8, atol=1e-8 * np.abs(op).max())
    assert np.allclose(vel, ov, rtol=1e-8, atol=1e-8 * np.abs(ov).max())
    # and it is not the beta = 0 run: the h history differs
    assert not np.allclose(gh, g["hist"][:200, 0], rtol=1e-6, atol=0)


def test_single_attempt_stage_combination_bit_identical(rng):
    # one fixed step without cosines: every stage combination bit-identical
    amu, m, ch, pos, vel = _rand_cloud(rng, 333)
    kw = dict(trap=0, coulomb=1, V0=10.0, d0=11e-3, Bz=7.0, delta=300.0)
    sf, of = _both_fields(**kw)
    with S.Engine(precision="fp64_exact", adaptive=False, use_graph=False, h0=3e-9) as e:
        e.add_cloud(pos, vel, m, ch, sf)
        e.step()
        gp, gv = e.state(0)
    c = O.Cloud(pos, vel, m, ch, of, h0=3e-9, adaptive=False, time_mode=O.TIME_FROZEN)
    c.step()
    op, ov, _ = c.state()
    assert np.array_equal(gp, op)
    assert np.array_equal(gv, ov)


def test_graph_and_eager_agree(rng):
    amu, m, ch, pos, vel = _rand_cloud(rng, 900, sigma=1e-3)
    sf, _ = _both_fields(trap=0, coulomb=1, V0=10.0, d0=11e-3, Bz=7.0, Ed=10.0, fd=940.0)
    res = []
    for use_graph in (False, True):
        with S.Engine(precision="mixed", use_graph=use_graph, rtol=1e-9, atol=1e-9,
                      time_mode=S.TIME_STAGE) as e:
            e.add_cloud(pos, vel, m, ch, sf)
            e.advance(2e-7)
            res.append((*e.state(0), e.status(0)))
    assert np.array_equal(res[0][0], res[1][0])
    assert np.array_equal(res[0][1], res[1][1])
    assert res[0][2]["attempts"] == res[1][2]["attempts"]


def test_graph_replays_match_eager_each_step(rng):
    """Every replay of the captured attempt equals the eager launch sequence,
    bit for bit, with eager work on the same stream between replays (a
    Coulomb evaluation, and a device re-sort every 3 attempts)."""
    amu, m, ch, pos, vel = _rand_cloud(rng, 700, sigma=1e-3)
    sf, _ = _both_fields(trap=0, coulomb=1, V0=10.0, d0=11e-3, Bz=7.0, Ed=10.0, fd=940.0)
    kw = dict(precision="mixed", rtol=1e-9, atol=1e-9, time_mode=S.TIME_STAGE, reorder_interval=3)
    with S.Engine(use_graph=True, **kw) as g, S.Engine(use_graph=False, **kw) as e:
        g.add_cloud(pos, vel, m, ch, sf)
        e.add_cloud(pos, vel, m, ch, sf)
        for k in range(12):
            g.step()
            e.step()
            fg = g.eval_coulomb(0, "mixed")
            fe = e.eval_coulomb(0, "mixed")
            assert np.array_equal(fg, fe), k
            (pg, vg), (pe, ve) = g.state(0), e.state(0)
            assert np.array_equal(pg, pe) and np.array_equal(vg, ve), k
            sg, se = g.status(0), e.status(0)
            assert (sg["lifetime"], sg["h"], sg["attempts"]) == (se["lifetime"], se["h"], se["attempts"]), k
            if k % 4 == 3:
                g.reorder()
                e.reorder()
        assert g.stats()["reorders"] > 0


def test_penning_physics_vs_oracle(rng):
    # stage-time mode with B, Penning quadrupole, dipole + quadrupole drive, friction
    amu, m, ch, pos, vel = _rand_cloud(rng, 200, sigma=5e-4, species=(133.0,), charges=(1,))
    nu_c, _, nu_m, _ = synth.penning_frequencies(133.0)
    kw = dict(trap=0, coulomb=1, V0=10.0, d0=11e-3, Bz=7.0, Ed=10.0, fd=nu_m, Uq=0.1,
              rq=11e-3, fq=nu_c, delta=1e3)
    sf, of = _both_fields(**kw)
    with S.Engine(precision="fp64_exact", rtol=1e-9, atol=1e-9, time_mode=S.TIME_STAGE) as e:
        e.add_cloud(pos, vel, m, ch, sf)
        for _ in range(40):
            e.step()
        gp, gv = e.state(0)
        st = e.status(0)
    c = O.Cloud(pos, vel, m, ch, of, h0=1e-9, rtol=1e-9, atol=1e-9, time_mode=O.TIME_STAGE)
    for _ in range(40):
        c.step()
    op, ov, sc = c.state()
    assert st["attempts"] == sc["attempts"]
    assert st["lifetime"] == pytest.approx(sc["lifetime"], rel=1e-10)
    assert np.allclose(gp, op, rtol=1e-9, atol=1e-9 * np.abs(op).max())
    assert np.allclose(gv, ov, rtol=1e-9, atol=1e-9 * np.abs(ov).max())


def test_ensemble_clouds_have_independent_controllers(rng):
    # uneven clouds, own start time, drive frequency and step-size history,
    # advanced together; each must follow its own CPU trajectory
    sizes, t0s, fds = [50, 200, 131], [0.0, 3e-7, 1.1e-6], [900.0, 2.0e5, 7.7e5]
    clouds = []
    with S.Engine(precision="fp64_exact", rtol=1e-9, atol=1e-9, time_mode=S.TIME_STAGE) as e:
        for n, t0, fd in zip(sizes, t0s, fds):
            amu, m, ch, pos, vel = _rand_cloud(rng, n, species=(133.0, 85.0), charges=(1,))
            kw = dict(trap=0, coulomb=1, V0=10.0, d0=11e-3, Bz=7.0, Ed=30.0, fd=fd)
            sf, of = _both_fields(**kw)
            e.add_cloud(pos, vel, m, ch, sf, lifetime=t0)
            clouds.append((pos, vel, m, ch, of, t0))
        e.advance(1.5e-7)
        got = [(*e.state(i), e.status(i)) for i in range(3)]
    for (pos, vel, m, ch, of, t0), (gp, gv, st) in zip(clouds, got):
        c = O.Cloud(pos, vel, m, ch, of, lifetime=t0, h0=1e-9, rtol=1e-9, atol=1e-9,
                    time_mode=O.TIME_STAGE)
        c.advance(t0 + 1.5e-7)
        op, ov, sc = c.state()
        assert st["attempts"] == sc["attempts"] and st["steps"] == sc["steps"]
        assert st["lifetime"] == pytest.approx(sc["lifetime"], rel=1e-12)
        assert st["lifetime"] >= t0 + 1.5e-7
        assert np.allclose(gp, op, rtol=1e-10, atol=1e-10 * np.abs(op).max())
        assert np.allclose(gv, ov, rtol=1e-10, atol=1e-10 * np.abs(ov).max())


def test_stage1_reuse_counts_and_exactness(rng):
    # Penning cloud in stage-time mode: every attempt after the first reuses
    # its stage-1 derivative (6 Coulomb evaluations), with results identical
    # to the oracle, which always evaluates 7 times
    amu, m, ch, pos, vel = _rand_cloud(rng, 150, species=(133.0,), charges=(1,))
    kw = dict(trap=0, coulomb=1, V0=10.0, d0=11e-3, Bz=7.0, Ed=20.0, fd=3e5)
    sf, of = _both_fields(**kw)
    with S.Engine(precision="fp64_exact", rtol=1e-9, atol=1e-9, time_mode=S.TIME_STAGE) as e:
        e.add_cloud(pos, vel, m, ch, sf)
        for _ in range(25):
            e.step()
        st = e.status(0)
        stats = e.stats()
        gp, _ = e.state(0)
    assert st["coulomb_evals"] == 7 + 6 * (st["attempts"] - 1)
    assert stats["pair_interactions"] == st["coulomb_evals"] * 150 * 149
    c = O.Cloud(pos, vel, m, ch, of, h0=1e-9, rtol=1e-9, atol=1e-9, time_mode=O.TIME_STAGE)
    for _ in range(25):
        c.step()
    op, _, sc = c.state()
    assert st["attempts"] == sc["attempts"]
    assert np.allclose(gp, op, rtol=1e-10, atol=1e-10 * np.abs(op).max())
    # frozen time with an RF field: the stage-7 derivative is not the next
    # stage 1 (different time), so only rejected attempts reuse it
    kw = dict(trap=1, Vrf=50.0, Vdc=5.0, kappa=0.001, r0=0.04, z0=0.04, freq_rf=5e6, coulomb=1)
    sf, _ = _both_fields(**kw)
    with S.Engine(precision="fp64_exact", adaptive=False, h0=1e-9, time_mode=S.TIME_FROZEN) as e:
        e.add_cloud(pos, vel, m, ch, sf)
        for _ in range(5):
            e.step()
        st = e.status(0)
    assert st["attempts"] == 5 and st["coulomb_evals"] == 35


def test_trap_ramp_on_device_matches_oracle(rng):
    # the Vrf/Vdc ramp of move_particles (bin/ion_fly.cpp:235-239), evaluated on
    # the device from the cloud's lifetime once per step
    amu, m, ch, pos, vel = _rand_cloud(rng, 40, species=(133.0,), charges=(1,))
    kw = dict(trap=1, Vrf=50.0, Vdc=5.0, kappa=0.001, r0=0.04, z0=0.04, freq_rf=5e6, coulomb=1,
              ramp=1, ramp_t0=0.0, ramp_t1=1e-6, newVrf=20.0, newVdc=9.0)
    sf, of = _both_fields(**kw)
    with S.Engine(precision="fp64_exact", adaptive=False, h0=1e-8, use_graph=True) as e:
        e.add_cloud(pos, vel, m, ch, sf)
        e.advance(3e-7)
        gp, gv = e.state(0)
        st = e.status(0)
    c = O.Cloud(pos, vel, m, ch, of, h0=1e-8, adaptive=False)
    k, _ = c.advance(3e-7)
    op, ov, sc = c.state()
    assert st["steps"] == k >= 30
    assert np.allclose(gp, op, rtol=1e-12, atol=1e-12 * np.abs(op).max())
    assert np.allclose(gv, ov, rtol=1e-11, atol=1e-11 * np.abs(ov).max())


def test_set_step_and_set_fields_between_runs(rng):
    amu, m, ch, pos, vel = _rand_cloud(rng, 64, species=(133.0,), charges=(1,))
    kw = dict(trap=0, coulomb=1, V0=10.0, d0=11e-3, Bz=7.0)
    sf, of = _both_fields(**kw)
    with S.Engine(precision="fp64_exact", adaptive=False, h0=1e-9, use_graph=True) as e:
        e.add_cloud(pos, vel, m, ch, sf)
        e.attempts(5)
        e.set_step(0, 2e-9)
        e.set_fields(0, S.default_fields(**dict(kw, delta=500.0)))
        e.attempts(5)
        e.synchronize()
        gp, gv = e.state(0)
        st = e.status(0)
    c = O.Cloud(pos, vel, m, ch, of, h0=1e-9, adaptive=False)
    for _ in range(5):
        c.step()
    of2 = O.fields(**dict(kw, delta=500.0))
    p1, v1, sc1 = c.state()
    c2 = O.Cloud(p1, v1, m, ch, of2, lifetime=sc1["lifetime"], h0=2e-9, adaptive=False)
    for _ in range(5):
        c2.step()
    op, ov, sc = c2.state()
    assert st["steps"] == 10 and st["lifetime"] == pytest.approx(sc["lifetime"], rel=1e-15)
    assert np.array_equal(gp, op) and np.array_equal(gv, ov)


def test_mixed_short_horizon_deviation(rng):
    # 100 fixed steps: mixed vs FP64 engine trajectories stay close
    w = synth.workload(2, n_override=4096)
    c = w.clouds[0]
    out = {}
    for p in ("fp64", "mixed"):
        with S.Engine(precision=p, adaptive=False, h0=2e-9, time_mode=S.TIME_STAGE) as e:
            e.add_cloud(c.pos, c.vel, c.mass, c.charge, c.fields)
            e.attempts(100)
            e.synchronize()
            out[p] = e.state(0)
    dev = np.abs(out["mixed"][0] - out["fp64"][0]).max()
    assert dev < 1e-9 * np.abs(out["fp64"][0]).max()


def test_coulomb_at_one_million_ions():
    """2^20 ions, beyond every benchmark set-up (index arithmetic of items,
    slices and partial sums at scale): the mixed sum, spatially ordered with
    local coordinates, against the oracle's FP64 rows on sampled targets."""
    n = 1 << 20
    c = synth.gaussian_cloud(n, 3e-3, 300.0, species=((133.0, 0.7), (85.0, 0.3)), seed=9)
    c.fields = synth.penning_fields()
    print(f"Coulomb at {n} ions", flush=True)
    with S.Engine(precision="mixed") as e:
        e.add_cloud(c.pos, c.vel, c.mass, c.charge, c.fields)
        e.set_launch(4, 4, 0, 1)
        f = e.eval_coulomb(0, "mixed")
    rows = np.random.Generator(np.random.PCG64(20)).choice(n, 48, replace=False).astype(np.int32)
    ref = O.coulomb_rows(c.pos, c.mass, c.charge, c.fields.coulomb_scale, rows)
    assert _rows_error(f[rows], ref) <= 1e-6
    assert np.all(np.isfinite(f))


@pytest.mark.parametrize("index,steps", [(3, 20), (5, 4)])
def test_full_size_short_horizon(index, steps):
    """The 65,536-ion and 262,144-ion set-ups at full size: fixed steps of the
    workload's mixed precision stay within 1e-11 of the cloud's extent from
    the FP64 engine (stage time, as the benchmark)."""
    w = synth.workload(index)
    c = w.clouds[0]
    print(f"full-size short horizon, workload {index}: {c.n} ions, {steps} steps", flush=True)
    out = {}
    for p in ("fp64", w.precision):
        with S.Engine(precision=p, adaptive=False, h0=1e-9, time_mode=S.TIME_STAGE) as e:
            e.add_cloud(c.pos, c.vel, c.mass, c.charge, c.fields)
            e.attempts(steps)
            e.synchronize()
            out[p] = e.state(0)
            assert e.status(0)["steps"] == steps
    scale = np.abs(out["fp64"][0]).max()
    assert np.abs(out[w.precision][0] - out["fp64"][0]).max() <= 1e-11 * scale
    vscale = np.abs(out["fp64"][1]).max()
    assert np.abs(out[w.precision][1] - out["fp64"][1]).max() <= 1e-8 * vscale


def test_advance_reaches_t_end_and_counts(rng):
    amu, m, ch, pos, vel = _rand_cloud(rng, 128, species=(133.0,), charges=(1,))
    sf, _ = _both_fields(trap=0, coulomb=1, V0=10.0, d0=11e-3, Bz=7.0)
    with S.Engine(precision="mixed", rtol=1e-8, atol=1e-8, time_mode=S.TIME_STAGE) as e:
        e.add_cloud(pos, vel, m, ch, sf, lifetime=1e-6)
        e.reset_stats()
        e.advance(3e-7)
        st = e.status(0)
        stats = e.stats()
    assert st["lifetime"] >= 1.3e-6 and st["active"] == 0
    assert stats["steps"] == st["steps"]
    assert stats["pair_interactions"] == st["coulomb_evals"] * 128 * 127
    assert 6 * st["attempts"] < st["coulomb_evals"] <= 7 * st["attempts"]


def test_underflow_is_reported():
    sf, _ = _both_fields(trap=1, Vrf=0.0, kappa=1.0, Vdc=5.0, z0=11e-3, r0=1.0)
    with S.Engine(precision="fp64_exact", rtol=0.0, atol=1e-300, h0=1e-6) as e:
        e.add_cloud([[1e-3, 0, 1e-3]], [[100.0, 0, 0]], O.mass_kg([133.0]), [1], sf)
        with pytest.raises(S.StepsizeUnderflow):
            e.step()


# ---------------------------------------------------------------------------
# cunbody1_force ABI
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("n", [1, 100, 1000, 4099])
def test_cunbody_abi_exact(rng, n):
    x = rng.normal(0, 1e-3, (n, 3))
    mj = np.ones(n)
    eps2 = float(np.float32(1e-23))   # bin/cunbody.cpp:11,66
    ai = S.cunbody1_force(x, mj, x.copy(), eps2, precision="fp64_exact")
    ref = O.cunbody_force(x, mj, x, eps2)
    assert np.array_equal(ai, ref)


@pytest.mark.parametrize("n", [2, 257, 3000])
def test_cunbody_abi_eps2_zero_masks_the_self_pair(rng, n):
    """eps2 = 0 with the reference's call (xj = pos2, xi = a copy of it,
    bin/cunbody.cpp:65-66): each ion meets itself at r = 0.  Every precision
    must stay finite and equal the oracle, which skips that pair."""
    x = rng.normal(0, 1e-3, (n, 3))
    mj = np.ones(n)
    ref = O.cunbody_force(x, mj, x.copy(), 0.0)
    assert np.isfinite(ref).all()
    for p, tol in (("fp64_exact", 0.0), ("fp64", 1e-13), ("mixed", 1e-6), ("mixed_kahan", 1e-6)):
        print(f"cunbody eps2=0 n={n} {p}", flush=True)
        ai = S.cunbody1_force(x, mj, x.copy(), 0.0, precision=p)
        assert np.isfinite(ai).all(), p
        if tol == 0.0:
            assert np.array_equal(ai, ref), p
        else:
            assert _rows_error(ai, ref) < tol, p


def test_cunbody_abi_fp64_and_mixed(rng):
    ni, nj = 3000, 5000
    xj = rng.normal(0, 1e-3, (nj, 3)); xi = rng.normal(0, 1e-3, (ni, 3))
    mj = rng.uniform(0.5, 2.0, nj)
    ref = O.cunbody_force(xj, mj, xi, 1e-20)
    a64 = S.cunbody1_force(xj, mj, xi, 1e-20, precision="fp64")
    amx = S.cunbody1_force(xj, mj, xi, 1e-20, precision="mixed")
    assert _rows_error(a64, ref) < 1e-13
    assert _rows_error(amx, ref) < 1e-6
    # unit weights (what the reference passes, bin/cunbody.cpp:34-35) take the q == 1 kernel
    ones = np.ones(nj)
    ref1 = O.cunbody_force(xj, ones, xi, 1e-20)
    assert _rows_error(S.cunbody1_force(xj, ones, xi, 1e-20, precision="mixed"), ref1) < 1e-6
    assert _rows_error(S.cunbody1_force(xj, ones, xi, 1e-20, precision="mixed_kahan"), ref1) < 1e-6


def test_cunbody_abi_empty_sets_and_changing_sizes(rng):
    # no sources: every sum is empty, in every precision
    xi = rng.normal(0, 1e-3, (10, 3))
    for p in ("fp64_exact", "fp64", "mixed", "mixed_kahan"):
        ai = S.cunbody1_force(np.zeros((0, 3)), np.zeros(0), xi, 1e-20, precision=p)
        assert np.array_equal(ai, np.zeros((10, 3))), p
    # no targets
    assert S.cunbody1_force(xi, np.ones(10), np.zeros((0, 3)), 1e-20).shape == (0, 3)
    # sizes that grow, shrink and grow again: the cached device buffers, work
    # lists and slice partials are reused or regrown between calls
    for ni, nj in ((100, 300), (5000, 7000), (50, 20), (9000, 300), (300, 12000), (64, 64)):
        xj = rng.normal(0, 1e-3, (nj, 3)); xs = rng.normal(0, 1e-3, (ni, 3))
        mj = rng.uniform(0.5, 2.0, nj)
        ref = O.cunbody_force(xj, mj, xs, 1e-20)
        assert _rows_error(S.cunbody1_force(xj, mj, xs, 1e-20, precision="fp64"), ref) < 1e-13, (ni, nj)
        assert _rows_error(S.cunbody1_force(xj, mj, xs, 1e-20, precision="mixed"), ref) < 1e-6, (ni, nj)


# ---------------------------------------------------------------------------
# NBODY ABI (initialize / simulation_Async / finalize)
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("n", [1, 64, 1000, 5003])
def test_nbody_abi(rng, n):
    pos32 = rng.normal(0, 1e-3, (n, 3)).astype(np.float32)
    acc = S.nbody_accel(pos32)
    x = pos32.astype(np.float64)
    ref = np.zeros((n, 3))
    for i in range(n):
        d = x - x[i]
        r2 = (d * d).sum(1) + float(np.float32(1e-7))
        ref[i] = (d / r2[:, None] ** 1.5).sum(0)
    if n == 1:
        assert np.all(acc == 0)
        return
    err = np.linalg.norm(acc - ref) / np.linalg.norm(ref)
    assert err < 1e-6, err


def test_add_cloud_rejects_non_finite_state(rng):
    amu, m, ch, pos, vel = _rand_cloud(rng, 10)
    sf, _ = _both_fields(trap=0, coulomb=1)
    for bad_pos in (True, False):
        p, v = pos.copy(), vel.copy()
        (p if bad_pos else v)[3, 1] = np.nan if bad_pos else np.inf
        with S.Engine(precision="mixed") as e:
            with pytest.raises(S.SimbucaError, match="non-finite"):
                e.add_cloud(p, v, m, ch, sf)
