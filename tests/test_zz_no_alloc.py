"""sbk_cunbody_force_static: cunbody1_force (bin/cunbody.h:11-17) with every
device buffer in the library's static scratch array, for callers that cannot
allocate.  Through the C ABI, against cunbody1_force itself on the same
inputs: bit-identical in FP64-exact (one slice, same per-target order), and
within the tolerances of the parity tests in the tiled precisions (the slice
count may differ, which reorders the FP64 sum over slices).  Two calls in a
row, the second smaller, so stale scratch contents would show; sizes beyond
the maximum are rejected with an error and leave the output alone."""
import numpy as np
import pytest

import oracle as O
import simbuca_amd as S

pytestmark = pytest.mark.gpu

MAX_IONS = 32768


def _call_static(xj, mj, xi, eps2, precision):
    L = S.lib()
    xj = np.ascontiguousarray(xj, dtype=np.float64)
    xi = np.ascontiguousarray(xi, dtype=np.float64)
    mj = np.ascontiguousarray(mj, dtype=np.float64)
    ai = np.full_like(xi, np.nan)
    old = L.sbk_cunbody_set_precision(S.PRECISIONS[precision])
    try:
        rc = L.sbk_cunbody_force_static(S._dp(xj), S._dp(mj), S._dp(xi), float(eps2), S._dp(ai), xi.shape[0],
                          xj.shape[0])
    finally:
        L.sbk_cunbody_set_precision(old)
    return rc, ai


def _rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


TOL = {"fp64_exact": 0.0, "fp64": 1e-13, "fp64_rsq": 1e-13, "mixed": 1e-6, "mixed_kahan": 1e-6}


@pytest.mark.parametrize("precision", ["fp64_exact", "fp64", "fp64_rsq", "mixed", "mixed_kahan"])
def test_static_matches_cunbody1_force_twice(precision):
    rng = np.random.Generator(np.random.PCG64(21))
    # the reference's call (xi a copy of xj, mj = 1), then a smaller call with
    # distinct target and source sets and weights
    cases = [(20000, 20000, True), (3000, 5000, False)]
    for ni, nj, same in cases:
        print(f"static cunbody {precision} ni={ni} nj={nj}", flush=True)
        xj = rng.normal(0, 1e-3, (nj, 3))
        xi = xj.copy() if same else rng.normal(0, 1e-3, (ni, 3))
        mj = np.ones(nj) if same else rng.uniform(0.5, 2.0, nj)
        eps2 = float(np.float32(1e-23)) if same else 1e-20
        rc, got = _call_static(xj, mj, xi, eps2, precision)
        assert rc == 0, S.lib().sbk_last_error()
        want = S.cunbody1_force(xj, mj, xi, eps2, precision=precision)
        assert np.isfinite(got).all()
        if TOL[precision] == 0.0:
            assert np.array_equal(got, want)
            assert np.array_equal(got, O.cunbody_force(xj, mj, xi, eps2))
        else:
            assert _rel(got, want) < TOL[precision], _rel(got, want)


def test_static_limits_and_empty_sets():
    rng = np.random.Generator(np.random.PCG64(22))
    x = rng.normal(0, 1e-3, (MAX_IONS + 1, 3))
    rc, ai = _call_static(x, np.ones(MAX_IONS + 1), x[:10].copy(), 1e-20, "fp64")
    assert rc == -1 and np.isnan(ai).all()
    assert b"at most 32768" in S.lib().sbk_last_error()
    rc, ai = _call_static(x[:10], np.ones(10), x, 1e-20, "fp64")
    assert rc == -1
    # the largest accepted size works, and equals the allocating entry point
    xm = x[:MAX_IONS]
    rc, got = _call_static(xm, np.ones(MAX_IONS), xm.copy(), 0.0, "mixed")
    assert rc == 0
    want = S.cunbody1_force(xm, np.ones(MAX_IONS), xm.copy(), 0.0, precision="mixed")
    assert _rel(got, want) < 1e-6
    # no sources: zeros; no targets: nothing
    rc, ai = _call_static(np.zeros((0, 3)), np.zeros(0), x[:7].copy(), 1e-20, "fp64")
    assert rc == 0 and np.array_equal(ai, np.zeros((7, 3)))
    rc, ai = _call_static(x[:7], np.ones(7), np.zeros((0, 3)), 1e-20, "fp64")
    assert rc == 0 and ai.shape == (0, 3)
