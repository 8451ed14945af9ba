"""Persistent attempts (persistent = 1: a batch of attempts as one
cooperative launch with grid barriers between the phases) against the
graph replay of the one-phase kernels, and the separate launches (no
graph) against both.

The persistent kernel calls the device functions of the one-phase kernels
(Coulomb work items, stage blocks, the controller, the commit) and keeps
every sum's order, so states, statuses and statistics must agree bit for
bit with the graph path of the same launch shape: positions, velocities, h,
errold, attempt and step counts, accepted_last.  The cases cover every precision, 1 and 2 targets per
thread, unit and mixed charges, local coordinates on and off, rejected
attempts (a first step far too long), clouds that finish in the middle of a
batch, a cloud with its Coulomb switch off, an empty cloud, a crossing
plane and fixed steps.
"""
import numpy as np
import pytest

import oracle as O
import simbuca_amd as S

pytestmark = pytest.mark.gpu

PENNING = dict(trap=1, Vrf=0.0, V0=10.0, d0=11e-3, Bz=7.0, coulomb=1)


def _cloud(rng, n, sigma=5e-4, unit=False, amu=133.0):
    pos = rng.normal(0, sigma, (n, 3))
    vel = rng.normal(0, 137.0, (n, 3))
    m = O.mass_kg(np.full(n, amu))
    ch = np.ones(n, np.int32) if unit else rng.choice(np.array([1, 1, 2, -1], np.int32), n)
    return pos, vel, m, ch.astype(np.int32)


PATHS = {"graph": dict(), "persistent": dict(persistent=1)}


def _run(clouds, fields, precision, tpt, path, local=-1, duration=3e-7, h0=4e-7,
         crossing=None, adaptive=True, batches=None, use_graph=True):
    """Advances every cloud by `duration` and returns states, statuses, stats
    and crossings."""
    with S.Engine(precision=precision, rtol=1e-9, atol=1e-9, h0=h0, time_mode=S.TIME_STAGE,
                  adaptive=adaptive, use_graph=use_graph, **PATHS[path]) as e:
        for (pos, vel, m, ch), f in zip(clouds, fields):
            e.add_cloud(pos, vel, m, ch, f)
        if precision != "fp64_exact":
            e.set_launch(tpt, 4, 0, local)
        if crossing is not None:
            e.set_crossing_plane(0, crossing)
        if batches is None:
            e.advance(duration)
        else:
            for i, k in enumerate(batches):
                e.attempts(k, open_ended=(i == 0))
            e.synchronize()
        states = [e.state(c) for c in range(len(clouds))]
        status = [e.status(c) for c in range(len(clouds))]
        stats = e.stats()
        xc = e.take_crossings() if crossing is not None else None
    return states, status, stats, xc


def _same(a, b):
    (sa, ta, za, xa), (sb, tb, zb, xb) = a, b
    for c, ((pa, va), (pb, vb)) in enumerate(zip(sa, sb)):
        assert np.array_equal(pa, pb), f"cloud {c} positions"
        assert np.array_equal(va, vb), f"cloud {c} velocities"
    for c, (u, v) in enumerate(zip(ta, tb)):
        assert u == v, (c, u, v)
    for k in ("attempts", "steps", "pair_interactions", "ion_steps", "reorders"):
        assert za[k] == zb[k], k
    if xa is not None:
        for k in xa:
            assert np.array_equal(xa[k], xb[k]), k


CASES = [("fp64_exact", 1, -1, False), ("fp64", 1, -1, False), ("fp64", 2, -1, False),
         ("fp64_rsq", 2, -1, False), ("mixed", 1, 0, False), ("mixed", 2, 1, False),
         ("mixed", 1, 1, True), ("mixed_kahan", 2, 1, True), ("mixed_kahan", 1, 0, False),
         ("fp64", 4, -1, False), ("mixed", 4, 1, True), ("mixed_kahan", 4, 0, False)]


def _used(path, stats):
    return (stats["persistent_launches"] > 0) == (path == "persistent")


@pytest.mark.parametrize("path,precision,tpt,local,unit", [("persistent",) + c for c in CASES])
def test_path_equals_graph_single_cloud(rng, path, precision, tpt, local, unit):
    clouds = [_cloud(rng, 700, unit=unit)]
    fields = [S.default_fields(**PENNING)]
    print(f"single cloud {path} {precision} tpt={tpt} local={local} unit={unit}", flush=True)
    g = _run(clouds, fields, precision, tpt, "graph", local)
    p = _run(clouds, fields, precision, tpt, path, local)
    assert _used("graph", g[2]) and _used(path, p[2])
    assert p[1][0]["steps"] > 10 and p[1][0]["attempts"] > p[1][0]["steps"]   # rejects happened
    _same(g, p)


@pytest.mark.parametrize("path,precision,tpt,local,unit", [("persistent",) + c for c in CASES[::2]])
def test_path_equals_graph_ensemble(rng, path, precision, tpt, local, unit):
    """Clouds of different sizes, species and fields: they take different
    numbers of attempts, so some finish in the middle of a batch; one cloud
    has no Coulomb term and one is empty."""
    sizes = [3001, 64, 3, 0, 700, 257]
    clouds = [_cloud(rng, n, sigma=s, unit=unit, amu=a)
              for n, s, a in zip(sizes, [1e-3, 5e-4, 2e-4, 1e-3, 8e-4, 6e-4],
                                 [133.0, 85.0, 39.0, 133.0, 133.0, 85.0])]
    fields = [S.default_fields(**dict(PENNING, Bz=b)) for b in (7.0, 5.0, 7.0, 7.0, 3.0, 7.0)]
    fields[5] = S.default_fields(**dict(PENNING, coulomb=0))
    print(f"ensemble {path} {precision} tpt={tpt} local={local} unit={unit}", flush=True)
    g = _run(clouds, fields, precision, tpt, "graph", local, crossing=0.0)
    p = _run(clouds, fields, precision, tpt, path, local, crossing=0.0)
    assert _used("graph", g[2]) and _used(path, p[2])
    steps = [st["steps"] for st in p[1]]
    assert len(set(steps[:3] + steps[4:])) > 1            # the clouds took different step counts
    assert p[3]["ion"].size > 0                            # ions crossed x = 0
    _same(g, p)


def test_paths_batches_and_fixed_steps(rng):
    """Open-ended batches of odd sizes (every attempt runs, none ends early)
    with fixed steps: persistent and separate launches against the graph."""
    clouds = [_cloud(rng, 500), _cloud(rng, 130)]
    fields = [S.default_fields(**PENNING)] * 2
    runs = {}
    for path, graph in (("graph", True), ("persistent", True), ("graph", False)):
        runs[path, graph] = _run(clouds, fields, "mixed", 1, path, 1, h0=2e-9, adaptive=False,
                                 batches=[1, 3, 7, 2], use_graph=graph)
    for key in runs:
        _same(runs["graph", True], runs[key])
    assert runs["graph", True][1][0]["attempts"] == 13 and runs["graph", True][1][0]["steps"] == 13
    assert runs["persistent", True][2]["persistent_launches"] == 4


def test_persistent_is_opt_in(rng):
    """The default replays the graph (faster on MI355X at every size
    measured); persistent must be 0 or 1."""
    clouds = [_cloud(rng, 500)]
    fields = [S.default_fields(**PENNING)]
    for use_graph in (True, False):
        with S.Engine(precision="mixed", use_graph=use_graph, h0=1e-9) as e:
            e.add_cloud(*clouds[0], fields[0])
            e.set_launch(1, 4, 0, -1)
            e.advance(2e-8)
            assert e.stats()["persistent_launches"] == 0
    for bad in (-1, 2):
        with pytest.raises(S.SimbucaError, match="persistent"):
            S.Engine(persistent=bad)


def test_persistent_rejects_shape_without_kernel(rng):
    clouds = [_cloud(rng, 300)]
    with S.Engine(precision="mixed", persistent=1) as e:
        e.add_cloud(*clouds[0], S.default_fields(**PENNING))
        e.set_launch(8, 4, 0, -1)   # 8 targets per thread: no persistent kernel
        with pytest.raises(S.SimbucaError, match="persistent"):
            e.advance(1e-8)


def test_persistent_barrier_timeout_is_reported(rng):
    """A spin limit of one poll makes every grid-barrier wait give up at
    once: every work-group leaves the kernel (the flag makes later barriers
    fail without waiting) and the host reports the time-out."""
    clouds = [_cloud(rng, 3001), _cloud(rng, 700)]
    fields = [S.default_fields(**PENNING)] * 2
    with S.Engine(precision="mixed", persistent=1, persistent_spin=1, h0=1e-9) as e:
        for c in clouds:
            e.add_cloud(*c, fields[0])
        e.set_launch(1, 4, 0, -1)
        with pytest.raises(S.SimbucaError, match="timed out"):
            e.advance(1e-7)
