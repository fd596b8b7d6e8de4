"""Exact transport LP values on the MI355X (transport.hip) against scipy
HiGHS golden values and the FP64 oracle's simplex.  Both sides are exact LP
solvers, so the optimum values agree to rounding: 1e-9 relative."""
import os

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden")


def test_lp_golden_highs(engine):
    print("transport golden", flush=True)
    g = np.load(os.path.join(GOLD, "lp_highs.npz"), allow_pickle=False)
    oc = ox = oy = 0
    for n, m, val in zip(g["n"], g["m"], g["val"]):
        n, m = int(n), int(m)
        Cm = g["C"][oc:oc + n * m]
        wX = g["wX"][ox:ox + n]
        wY = g["wY"][oy:oy + m]
        oc += n * m
        ox += n
        oy += m
        v = engine.transport_batch(Cm[None], wX[None], wY[None])[0]
        assert abs(v - val) <= 1e-9 * max(1.0, abs(val)), (n, m, v, val)


def test_lp_golden_highs_serving_shapes(engine):
    """HiGHS values at the serving shapes (48 x 200 ... 48 x 500, 64 x 500),
    one LP per launch and all of one shape in a batch."""
    from cases import lp_large_cases
    print("transport golden, serving shapes", flush=True)
    for n, m, Cm, wX, wY, val in lp_large_cases(np.load(os.path.join(GOLD, "lp_highs_large.npz"),
                                                        allow_pickle=False)):
        v = engine.transport_batch(Cm[None], wX[None], wY[None])[0]
        assert abs(v - val) <= 1e-9 * max(1.0, abs(val)), (n, m, v, val)


@pytest.mark.parametrize("n,m,nb", [(6, 12, 64), (40, 80, 32), (64, 200, 8), (1, 7, 4), (9, 1, 4),
                                    (48, 300, 6), (48, 500, 4)])
def test_lp_batch_vs_oracle(engine, n, m, nb):
    print(f"transport batch n={n} m={m} nb={nb}", flush=True)
    rng = np.random.default_rng(n * 1000 + m)
    d = 64 if m >= 300 else 2  # the serving workload's d = 64 at its largest LPs
    A = rng.standard_normal((nb, n, d))
    B = rng.standard_normal((nb, m, d))
    C = ((A[:, :, None, :] - B[:, None, :, :]) ** 2).sum(-1)  # (nb, n, m)
    Cf = np.ascontiguousarray(C.transpose(0, 2, 1)).reshape(nb, -1)  # [i + j*n]
    wX = rng.dirichlet(np.ones(n), size=nb)
    wY = rng.dirichlet(np.ones(m), size=nb)
    got, piv = engine.transport_batch(Cf, wX, wY, pivots=True)
    for b in range(nb):
        ref = O.transport(Cf[b], wX[b], wY[b])
        assert abs(got[b] - ref) <= 1e-10 * max(1.0, abs(ref)), (b, got[b], ref)


def test_lp_degenerate_grid_costs(engine):
    """Integer grid supports (many equal costs, degenerate pivots) as in the
    16x16 image workload."""
    print("transport degenerate grid", flush=True)
    rng = np.random.default_rng(7)
    nb, n, m = 16, 40, 60
    Cs, wXs, wYs = [], [], []
    for _ in range(nb):
        a = rng.integers(0, 16, (n, 2)).astype(float)
        b = rng.integers(0, 16, (m, 2)).astype(float)
        Cs.append(((a[:, None] - b[None]) ** 2).sum(-1).T.reshape(-1))
        wXs.append(np.full(n, 1.0 / n))
        wYs.append(np.full(m, 1.0 / m))
    got = engine.transport_batch(np.array(Cs), np.array(wXs), np.array(wYs))
    for b in range(nb):
        ref = O.transport(Cs[b], wXs[b], wYs[b])
        assert abs(got[b] - ref) <= 1e-10 * max(1.0, abs(ref))


@pytest.mark.parametrize("n,m,nb", [(40, 50, 8), (48, 500, 2), (3, 4, 4)])
def test_lp_phase_profile(engine, n, m, nb):
    """d2mi_transport_profile: the same values and pivots as the plain batch,
    every phase time non-negative, the pivot-loop phases present whenever an
    LP pivoted, and the total within the call's wall time."""
    import time
    print(f"transport profile n={n} m={m} nb={nb}", flush=True)
    rng = np.random.default_rng(7 + n + m)
    A = rng.standard_normal((nb, n, 8))
    B = rng.standard_normal((nb, m, 8))
    C = ((A[:, :, None, :] - B[:, None, :, :]) ** 2).sum(-1)
    Cf = np.ascontiguousarray(C.transpose(0, 2, 1)).reshape(nb, -1)
    wX = rng.dirichlet(np.ones(n), size=nb)
    wY = rng.dirichlet(np.ones(m), size=nb)
    v0, p0 = engine.transport_batch(Cf, wX, wY, pivots=True)
    t0 = time.perf_counter()
    v1, p1, ph = engine.transport_profile(Cf, wX, wY)
    wall_us = (time.perf_counter() - t0) * 1e6
    assert np.array_equal(v0, v1) and np.array_equal(p0, p1)
    assert ph.shape == (nb, 12) and (ph >= 0).all()
    assert (ph[:, :8].sum(1) > 0).all() and (ph[:, :8].sum(1) < wall_us).all()
    assert (ph[:, 8:].sum(1) <= ph[:, 2] + 1.0).all()  # Vogel's steps lie inside the start
    assert (ph[p1 > 0][:, 4:7].sum(1) > 0).all()
