"""The allocation-free flux-table entry point against the allocating one:
identical tables on two calls in a row (the second with fewer facets, so the
device array keeps the first call's tail), and a clear error above the
array's maximum.  Runs alone in a fresh process and after the other tests."""
import numpy as np
import pytest

import puncpp_amd as P
from puncpp_amd import _native as N

pytestmark = pytest.mark.gpu

MAX_FACETS = 131072


def _tables(fn, normal, area, kind, vth, vd, k, a, rng_sp, n_iter):
    nf = area.shape[0]
    num, pmax = np.full(nf, np.nan), np.full(nf, np.nan)
    vd = np.ascontiguousarray(vd, dtype=np.float64)
    N.check(fn(nf, N.dptr(normal), N.dptr(area), kind, vth, N.dptr(vd), k, a, 7.0, 3, rng_sp, n_iter,
               N.dptr(num), N.dptr(pmax)))
    return num, pmax


def test_noalloc_flux_tables_match(gpu):
    print("[case] noalloc_flux_tables", flush=True)
    L = N.lib()
    hm = P.box_mesh(16, [-2.0] * 3, [2.0] * 3, sphere_r=0.7)
    m = P.Mesh(hm)
    ef = m.exterior_facet_arrays()
    normal = np.ascontiguousarray(ef["normal"])
    area = np.ascontiguousarray(ef["area"])
    nf = area.shape[0]
    # KappaCairns has neither table in closed form: the Monte Carlo fills both
    cases = [(3, 4.0, 0.3, [0.0, 0.0, 0.4], 0, nf), (2, 0.0, 0.5, [0.0, 0.2, 0.4], 1, nf // 3)]
    for kind, k, a, vd, rng_sp, n in cases:
        nrm, ar = np.ascontiguousarray(normal[:n]), np.ascontiguousarray(area[:n])
        ref = _tables(L.punc_gpu_vdf_flux_facets, nrm, ar, kind, 1.0, vd, k, a, rng_sp, 20000)
        got = _tables(L.punc_gpu_vdf_flux_facets_noalloc, nrm, ar, kind, 1.0, vd, k, a, rng_sp, 20000)
        assert np.isfinite(got[0]).all() and np.isfinite(got[1]).all()
        assert (got[0] > 0).all() and (got[1] > 0).all()
        np.testing.assert_array_equal(got[0], ref[0])
        np.testing.assert_array_equal(got[1], ref[1])


def test_noalloc_flux_tables_reject_above_maximum(gpu):
    print("[case] noalloc_flux_reject", flush=True)
    L = N.lib()
    n = MAX_FACETS + 1
    normal = np.zeros((n, 3))
    normal[:, 2] = 1.0
    area = np.ones(n)
    num, pmax = np.empty(n), np.empty(n)
    vd = np.zeros(3)
    rc = L.punc_gpu_vdf_flux_facets_noalloc(n, N.dptr(normal), N.dptr(area), 3, 1.0, N.dptr(vd), 4.0, 0.3, 7.0,
                                            0, 0, 1000, N.dptr(num), N.dptr(pmax))
    assert rc != 0 and "exceed the scratch maximum" in L.punc_gpu_last_error().decode()
