"""pdist2 on the matrix cores against the FP64 oracle (_dpdist2,
blas_like64.c:199-208).  FP32 products: relative error within 1e-5 of the
FP64 table (FP32 rounding of |x|^2 + |y|^2 - 2 x.y, with near-zero entries
recomputed by direct differences); BF16 products: within 2e-2."""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


@pytest.mark.parametrize("n,m,d", [(300, 1000, 300), (77, 129, 13), (1, 5, 1), (200, 64, 64)])
def test_pdist2_mfma_fp32(engine, n, m, d):
    say(f"pdist2 mfma fp32 {n}x{m}x{d}")
    rng = np.random.default_rng(n + m + d)
    A = rng.standard_normal((n, d))
    B = rng.standard_normal((m, d))
    A /= np.linalg.norm(A, axis=1, keepdims=True)
    B /= np.linalg.norm(B, axis=1, keepdims=True)
    # near duplicates: cancellation-prone entries
    k = min(n, m) // 3
    B[:k] = A[:k] + 1e-4 * rng.standard_normal((k, d))
    ref = O.pdist2(A, B)
    got = engine.pdist2_mfma(A, B)
    scale = np.maximum(ref, 1e-3)
    assert np.max(np.abs(got - ref) / scale) < 1e-5
    assert (got >= 0).all()


def test_pdist2_mfma_bf16(engine):
    say("pdist2 mfma bf16")
    rng = np.random.default_rng(3)
    A = rng.standard_normal((256, 300))
    B = rng.standard_normal((512, 300))
    A /= np.linalg.norm(A, axis=1, keepdims=True)
    B /= np.linalg.norm(B, axis=1, keepdims=True)
    ref = O.pdist2(A, B)
    got = engine.pdist2_mfma(A, B, bf16=True)
    assert np.max(np.abs(got - ref) / np.maximum(ref, 1e-2)) < 2e-2
