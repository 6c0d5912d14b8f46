"""Kernels of the learned surrogate against numpy: BF16 MFMA GEMM (all operand forms),
segment sums, gathers, Adam and the masked MSE."""
import numpy as np
import pytest

from uph_amd.devbuf import DevBuf, check, from_bf16, lib, to_bf16

pytestmark = pytest.mark.gpu


def _say(msg):
    print(f"[gpu-gnn] {msg}", flush=True)


def _gemm(M, N, K, A, lda, ta, B, ldb, out_bf16=0, act=0, splits=1, accumulate=False, C0=None,
          bias=None):
    a = DevBuf.of(to_bf16(A))
    b = DevBuf.of(to_bf16(B))
    ldc = N
    c = DevBuf(M * ldc * 4)
    if C0 is not None:
        c.upload(C0.astype(np.float32))
    work = DevBuf(M * N * 4 * max(splits, 1)) if (splits > 1 or accumulate) else None
    bb = DevBuf.of(bias.astype(np.float32)) if bias is not None else None
    check(lib().uph_gemm_bf16(M, N, K, a.ptr, lda, ta, b.ptr, ldb, bb.ptr if bb else None, c.ptr, ldc,
                              out_bf16, act, splits, work.ptr if work else None, int(accumulate),
                              None))
    check(lib().uph_dev_sync())
    if out_bf16:
        return from_bf16(c.download((M, ldc), np.uint16))
    return c.download((M, ldc), np.float32)


def _ref(A, B):
    return from_bf16(to_bf16(A)).astype(np.float64) @ from_bf16(to_bf16(B)).astype(np.float64)


@pytest.mark.parametrize("M,N,K", [(16, 16, 32), (64, 64, 32), (1000, 136, 136), (77, 8, 16)])
def test_gemm_plain(M, N, K):
    rng = np.random.default_rng(M + N + K)
    A = rng.normal(size=(M, K)).astype(np.float32)
    B = rng.normal(size=(K, N)).astype(np.float32)
    _say(f"gemm {M}x{N}x{K}")
    C = _gemm(M, N, K, A, K, 0, B, N)
    R = _ref(A, B)
    assert np.abs(C - R).max() < 1e-4 * np.abs(R).max() + 1e-5


def test_gemm_bias_relu_bf16_out():
    rng = np.random.default_rng(3)
    M, N, K = 300, 128, 136
    A = rng.normal(size=(M, K)).astype(np.float32)
    B = rng.normal(size=(K, N)).astype(np.float32)
    bias = rng.normal(size=N).astype(np.float32)
    C = _gemm(M, N, K, A, K, 0, B, N, out_bf16=1, act=1, bias=bias)
    R = np.maximum(_ref(A, B) + bias, 0)
    assert np.abs(C - R).max() < 1e-2 * np.abs(R).max()


@pytest.mark.parametrize("splits", [1, 7])
def test_gemm_trans_a_split_k(splits):
    """dW = X^T dY with a long reduction: X is rows x 136, dY rows x 128."""
    rng = np.random.default_rng(5)
    rows, F, N = 5000, 136, 128
    X = rng.normal(size=(rows, F)).astype(np.float32)
    D = rng.normal(size=(rows, N)).astype(np.float32)
    _say(f"gemm trans_a splits={splits}")
    C = _gemm(F, N, rows, X, F, 1, D, N, splits=splits)
    R = _ref(X.T.copy(), D)
    assert np.abs(C - R).max() < 1e-4 * np.abs(R).max()


def test_gemm_accumulate():
    rng = np.random.default_rng(6)
    M, N, K = 200, 64, 64
    A = rng.normal(size=(M, K)).astype(np.float32)
    B = rng.normal(size=(K, N)).astype(np.float32)
    C0 = rng.normal(size=(M, N)).astype(np.float32)
    C = _gemm(M, N, K, A, K, 0, B, N, accumulate=True, C0=C0)
    R = _ref(A, B) + C0
    assert np.abs(C - R).max() < 1e-4 * np.abs(R).max()


def test_gemm_rejects_bad_shapes():
    a = DevBuf(64)
    with pytest.raises(RuntimeError):
        check(lib().uph_gemm_bf16(4, 4, 4, a.ptr, 4, 0, a.ptr, 4, None, a.ptr, 4, 0, 0, 1, None, 0,
                                  None))


def test_segsum_and_gather():
    rng = np.random.default_rng(7)
    nseg, F, E = 50, 16, 400
    cuts = np.sort(rng.integers(0, E, nseg - 1))
    off = np.concatenate([[0], cuts, [E]]).astype(np.int32)
    idx = rng.permutation(E).astype(np.int32)
    x = rng.normal(size=(E, F)).astype(np.float32)
    dx, doff, didx = DevBuf.of(x), DevBuf.of(off), DevBuf.of(idx)
    out = DevBuf(nseg * F * 4)
    check(lib().uph_segsum(nseg, F, doff.ptr, didx.ptr, dx.ptr, F, 0, out.ptr, F, 0, None))
    check(lib().uph_dev_sync())
    got = out.download((nseg, F), np.float32)
    ref = np.stack([x[idx[off[i]:off[i + 1]]].sum(0) for i in range(nseg)])
    assert np.allclose(got, ref, atol=1e-5)
    rows = rng.integers(0, E, 123).astype(np.int32)
    y = DevBuf(123 * F * 4)
    drows = DevBuf.of(rows)            # keep every device buffer alive across the launch
    check(lib().uph_gather_rows(123, F, drows.ptr, dx.ptr, F, y.ptr, F, 4, None))
    assert np.array_equal(y.download((123, F), np.float32), x[rows])


def test_adam_and_mse():
    rng = np.random.default_rng(8)
    n = 1000
    w, g = rng.normal(size=n).astype(np.float32), rng.normal(size=n).astype(np.float32)
    m, v = np.zeros(n, np.float32), np.zeros(n, np.float32)
    dw, dg, dm, dv = DevBuf.of(w), DevBuf.of(g), DevBuf.of(m), DevBuf.of(v)
    wb = DevBuf(2 * n)
    check(lib().uph_adam(n, dw.ptr, dg.ptr, dm.ptr, dv.ptr, wb.ptr, 1e-3, 0.9, 0.999, 1e-8, 1, None))
    check(lib().uph_dev_sync())
    mr = 0.1 * g
    vr = 0.001 * g * g
    wr = w - 1e-3 * (mr / 0.1) / (np.sqrt(vr / 0.001) + 1e-8)
    assert np.allclose(dw.download(n, np.float32), wr, rtol=1e-5, atol=1e-7)
    assert np.allclose(from_bf16(wb.download(n, np.uint16)), wr, rtol=1e-2)
    M, Cc = 300, 2
    y, t = rng.normal(size=(M, Cc)).astype(np.float32), rng.normal(size=(M, Cc)).astype(np.float32)
    mask = (rng.random(M) < 0.7).astype(np.int32)
    cnt = int(mask.sum()) * Cc
    dy, lp = DevBuf(M * Cc * 4), DevBuf(((M * Cc + 255) // 256) * 4)
    yd, td, md = DevBuf.of(y), DevBuf.of(t), DevBuf.of(mask)
    check(lib().uph_mse(M, Cc, yd.ptr, Cc, td.ptr, Cc, md.ptr, cnt, dy.ptr, lp.ptr, None))
    check(lib().uph_dev_sync())
    loss = lp.download((M * Cc + 255) // 256, np.float32).sum()
    d = (y - t) * mask[:, None]
    assert loss == pytest.approx((d * d).sum() / cnt, rel=1e-5)
    assert np.allclose(dy.download((M, Cc), np.float32), 2 * d / cnt, atol=1e-7)


@pytest.mark.parametrize("ta", [0, 1])
def test_gemm_b_given_transposed(ta):
    """trans bit 1: B passed as its N x K transpose (with and without A^T)."""
    rng = np.random.default_rng(21 + ta)
    M, N, K = 200, 136, 144
    A = rng.normal(size=(M, K)).astype(np.float32)
    B = rng.normal(size=(K, N)).astype(np.float32)
    Ain = np.ascontiguousarray(A.T) if ta else A
    lda = M if ta else K
    C = _gemm(M, N, K, Ain, lda, ta | 2, np.ascontiguousarray(B.T), K)
    R = _ref(A, B)
    assert np.abs(C - R).max() < 1e-4 * np.abs(R).max() + 1e-5
