"""pdist2 on the matrix cores against the FP64 oracle (_dpdist2,
blas_like64.c:199-208).  FP32 products: relative error within 1e-5 of the
FP64 table (FP32 rounding of |x|^2 + |y|^2 - 2 x.y, with near-zero entries
recomputed by direct differences); BF16 products: the exact squared distance
of the BF16-rounded rows to FP32 accumulation error."""
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


def bf16_round(x):
    """FP64 -> FP32 -> BF16 with round-to-nearest-even, back as FP64."""
    b = np.ascontiguousarray(x, np.float32).view(np.uint32).astype(np.uint64)
    b = (b + 0x7FFF + ((b >> 16) & 1)) & 0xFFFF0000
    return b.astype(np.uint32).view(np.float32).astype(np.float64)


@pytest.mark.parametrize("n,m,d", [(256, 512, 300), (77, 129, 13), (1, 5, 1), (200, 64, 64),
                                   (130, 1300, 40)])
def test_pdist2_mfma_bf16(engine, n, m, d):
    """BF16 products: every entry is |x|^2 + |y|^2 - 2 bf16(x).bf16(y) with
    the norms of the FP32 rows and the products accumulated in FP32, so
    against that expression in FP64 it is exact to FP32 accumulation error,
    1e-5 of |x|^2 + |y|^2; entries small next to that (near-duplicate rows)
    are recomputed from the FP32 rows by direct differences and agree with
    the FP64 oracle on the unrounded rows to 1e-5 relative."""
    say(f"pdist2 mfma bf16 {n}x{m}x{d}")
    rng = np.random.default_rng(3 + n + d)
    A = rng.standard_normal((n, d))
    B = rng.standard_normal((m, d))
    A /= np.linalg.norm(A, axis=1, keepdims=True)
    B /= np.linalg.norm(B, axis=1, keepdims=True)
    k = min(n, m) // 3
    B[:k] = A[:k] + 1e-4 * rng.standard_normal((k, d))
    got = engine.pdist2_mfma(A, B, bf16=True).reshape(m, n)
    A32 = A.astype(np.float32).astype(np.float64)
    B32 = B.astype(np.float32).astype(np.float64)
    A16, B16 = bf16_round(A), bf16_round(B)
    norms = (A32 ** 2).sum(1)[None, :] + (B32 ** 2).sum(1)[:, None]
    ref16 = norms - 2.0 * (B16 @ A16.T)  # [i, j]: |A_j|^2 + |B_i|^2 - 2 A16_j . B16_i
    ref = O.pdist2(A32, B32).reshape(m, n)
    far = ref16 > 0.02 * norms
    near = ref16 < 0.005 * norms
    assert far.any()
    assert np.max(np.abs(got - ref16)[far] / norms[far]) < 1e-5
    if near.any():
        assert np.max(np.abs(got - ref)[near] / np.maximum(ref[near], 1e-7)) < 1e-5
    assert (got >= 0).all()


@pytest.mark.parametrize("d", [13, 40, 96, 300])
def test_pdist2_bf16_wide_equals_narrow(engine, d):
    """Tables of >= 192 columns take the kernel that copies its BF16 slices
    straight into three LDS stages (rotated 16-byte chunks); narrower ones
    the register-staged kernel.  Both multiply the same BF16 operands in the
    same order of the shared dimension, so a 256-column table must equal, bit
    for bit, the two 128-column tables of its halves; d covers 1, 2, 3 and 10
    slices of the three-stage pipeline."""
    say(f"pdist2 bf16 wide vs narrow, d={d}")
    rng = np.random.default_rng(11 + d)
    m, n = 777, 256
    A = rng.standard_normal((n, d))
    B = rng.standard_normal((m, d))
    A /= np.linalg.norm(A, axis=1, keepdims=True)
    B /= np.linalg.norm(B, axis=1, keepdims=True)
    B[:40] = A[:40] + 1e-4 * rng.standard_normal((40, d))  # recomputed entries too
    wide = engine.pdist2_mfma(A, B, bf16=True).reshape(m, n)
    lo = engine.pdist2_mfma(A[:128], B, bf16=True).reshape(m, 128)
    hi = engine.pdist2_mfma(A[128:], B, bf16=True).reshape(m, 128)
    np.testing.assert_array_equal(wide[:, :128], lo)
    np.testing.assert_array_equal(wide[:, 128:], hi)


@pytest.mark.parametrize("bf16", [False, True])
def test_pdist2_concentrated_recomputes(engine, bf16):
    """Entries that lose digits to cancellation are recomputed by direct
    differences of the FP32 rows.  Many centroid rows equal (or within 1e-4
    of) the same few vocabulary rows, as popular words make them, which puts
    hundreds of recomputes in a few tiles; they must equal the FP64
    distances of the FP32 rows to FP32 rounding, and the rest of the table
    must be untouched by them."""
    say(f"pdist2 concentrated recomputes, bf16={bf16}")
    rng = np.random.default_rng(29)
    V, Q, d = 3000, 320, 300
    vocab = rng.standard_normal((V, d))
    vocab /= np.linalg.norm(vocab, axis=1, keepdims=True)
    cs = vocab[rng.integers(0, 8, Q)] + np.where(rng.random((Q, 1)) < 0.5, 0.0, 1e-4) * \
        rng.standard_normal((Q, d))
    T = engine.pdist2_mfma(cs, vocab, bf16=bf16).reshape(V, Q)
    ref = O.pdist2(cs.astype(np.float32).astype(np.float64),
                   vocab.astype(np.float32).astype(np.float64)).reshape(V, Q)
    small = ref < 1e-3
    assert small.sum() >= Q // 2
    assert np.max(np.abs(T[small] - ref[small])) < 1e-6
    # the rest: FP32 rounding of |x|^2 + |y|^2 - 2 x.y (|x|^2 + |y|^2 = 2), or the
    # BF16 products' rounding (test_pdist2_mfma_bf16 pins that error exactly)
    assert np.max(np.abs(T[~small] - ref[~small])) < (1e-2 if bf16 else 1e-5)
