"""The CPU-runnable benchmark configuration on the GPU, as stated: 500 singly
charged 133 u ions, Gaussian sigma = 0.5 mm, 300 K Maxwell-Boltzmann, seed 1,
B = 7 T, ideal quadrupole V0 = 10 V with d0 = 11 mm, direct-sum Coulomb,
Dormand-Prince 5(4) adaptive with rtol = atol = 1e-9 (the reference hard-codes
1e-7, bin/ode.cpp:515-516), stage-time fields, 100 us (simbuca_amd.synth
workload 1).

* The first 200 adaptive steps of the FP64-exact engine against the oracle's
  restatement of step() (bin/ode.cpp:459-494) driven as move_particles does
  (bin/ion_fly.cpp:83-101): the same accept/reject sequence and the same
  states to 1e-9 (the only difference is the GPU's fixed-order sum of the
  error norm, which moves h in its last bits).
* The whole 100 us in the plain-FP64 and the mixed engine: the cloud is
  conservative and axisymmetric, so total energy and canonical L_z are
  conserved; their relative drifts over the ~11,000 steps bound the
  integrator's error (the measured drifts are about -5e-8 and 2e-10).
"""
import numpy as np
import pytest

import oracle as O
import simbuca_amd as S
from simbuca_amd import synth

pytestmark = pytest.mark.gpu


def _workload():
    w = synth.workload(1)
    assert len(w.clouds) == 1 and w.clouds[0].n == 500
    assert (w.rtol, w.atol, w.t_sim, w.time_mode) == (1e-9, 1e-9, 100e-6, S.TIME_STAGE)
    c = w.clouds[0]
    f = c.fields
    assert (f.Bz, f.V0, f.d0, f.coulomb, f.trap) == (7.0, 10.0, 11e-3, 1, 0)
    return w, c


def test_workload1_first_200_steps_match_the_oracle():
    w, c = _workload()
    print("workload 1: 200 adaptive steps, fp64_exact vs oracle", flush=True)
    of = O.fields(trap=0, Bz=7.0, V0=10.0, d0=11e-3, coulomb=1, coulomb_scale=1.0)
    oc = O.Cloud(c.pos, c.vel, c.mass, c.charge, of, h0=w.h0, rtol=w.rtol, atol=w.atol,
                 adaptive=True, time_mode=O.TIME_STAGE)
    with S.Engine(precision="fp64_exact", time_mode=w.time_mode, rtol=w.rtol, atol=w.atol,
                  h0=w.h0) as e:
        e.add_cloud(c.pos, c.vel, c.mass, c.charge, c.fields)
        gh, oh = [], []
        for _ in range(200):
            gh.append(e.status(0)["h"])
            oh.append(oc.state()[2]["h"])
            e.step()
            oc.step()
        gp, gv = e.state(0)
        st = e.status(0)
    op, ov, sc = oc.state()
    assert st["steps"] == sc["steps"] == 200
    assert st["attempts"] == sc["attempts"]
    assert np.allclose(gh, oh, rtol=1e-9, atol=0)
    assert st["lifetime"] == pytest.approx(sc["lifetime"], rel=1e-9)
    assert np.allclose(gp, op, rtol=1e-9, atol=1e-9 * np.abs(op).max())
    assert np.allclose(gv, ov, rtol=1e-9, atol=1e-9 * np.abs(ov).max())


@pytest.mark.parametrize("precision", ["fp64", "fp64_rsq", "mixed"])
def test_workload1_full_100us_conserves_energy_and_lz(precision):
    w, c = _workload()
    print(f"workload 1: 100 us, {precision}", flush=True)
    e0, l0 = synth.cloud_invariants(c, c.pos, c.vel)
    with S.Engine(precision=precision, time_mode=w.time_mode, rtol=w.rtol, atol=w.atol,
                  h0=w.h0) as e:
        e.add_cloud(c.pos, c.vel, c.mass, c.charge, c.fields)
        e.advance(w.t_sim)
        pos, vel = e.state(0)
        st = e.status(0)
    assert st["lifetime"] >= w.t_sim and st["error"] == 0
    assert 5000 < st["steps"] < 50000, st["steps"]
    assert np.isfinite(pos).all() and np.isfinite(vel).all()
    e1, l1 = synth.cloud_invariants(c, pos, vel)
    de, dl = (e1 - e0) / abs(e0), (l1 - l0) / abs(l0)
    print(f"workload 1 {precision}: {st['steps']} steps, energy drift {de:.3e}, "
          f"L_z drift {dl:.3e}", flush=True)
    assert abs(de) < 1e-6, de
    assert abs(dl) < 1e-8, dl
