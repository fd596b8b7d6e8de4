"""d2mi_transport_batch_noalloc, the LP batch that allocates nothing (its
staging lives in a fixed per-GPU scratch array), against the allocating
d2mi_transport_batch through the C ABI: two calls in a row, the second with
a smaller batch of smaller LPs (so it runs over what the first left in the
array), bit-identical values and pivot counts; a batch beyond the array is
refused with a clear error."""
import numpy as np
import pytest

from d2_kmeans_amd import _lib
from d2_kmeans_amd._lib import D2MIError, check, lib, ptr

pytestmark = pytest.mark.gpu


def _batch(rng, n, m, nb, d=3):
    A = rng.standard_normal((nb, n, d))
    B = rng.standard_normal((nb, m, d))
    Cm = ((A[:, :, None, :] - B[:, None, :, :]) ** 2).sum(-1)
    Cf = np.ascontiguousarray(Cm.transpose(0, 2, 1)).reshape(nb, -1)  # [i + j*n]
    return Cf, rng.dirichlet(np.ones(n), size=nb), rng.dirichlet(np.ones(m), size=nb)


def _noalloc(engine, Cf, wX, wY):
    nb, n = wX.shape
    m = wY.shape[1]
    out = np.full(nb, np.nan)
    piv = np.full(nb, -1, np.int32)
    check(lib().d2mi_transport_batch_noalloc(engine._h, n, m, nb, ptr(Cf), ptr(wX), ptr(wY),
                                             ptr(out), ptr(piv, _lib._ip)))
    return out, piv


def test_noalloc_lp_batch_matches_allocating_entry_point(engine):
    rng = np.random.default_rng(31)
    for n, m, nb in ((40, 90, 48), (9, 17, 5)):
        print(f"no-alloc LP batch n={n} m={m} nb={nb}", flush=True)
        Cf, wX, wY = _batch(rng, n, m, nb)
        ref, ref_piv = engine.transport_batch(Cf, wX, wY, pivots=True)
        got, got_piv = _noalloc(engine, Cf, wX, wY)
        np.testing.assert_array_equal(got, ref)
        np.testing.assert_array_equal(got_piv, ref_piv)


def test_noalloc_rejects_batches_beyond_its_scratch(engine):
    print("no-alloc oversize", flush=True)
    n, m = 48, 500
    nb = (32 << 20) // (8 * n * m) + 1
    Cf = np.zeros((nb, n * m))
    wX = np.full((nb, n), 1.0 / n)
    wY = np.full((nb, m), 1.0 / m)
    out = np.empty(nb)
    rc = lib().d2mi_transport_batch_noalloc(engine._h, n, m, nb, ptr(Cf), ptr(wX), ptr(wY),
                                            ptr(out), None)
    assert rc == -1  # D2MI_EINVAL
    assert b"scratch" in lib().d2mi_last_error()
    with pytest.raises(D2MIError):
        check(rc)
