"""Surrogate network kernels (include/cmf_mlp.h) against the CPU oracle
(oracle/mlp_oracle.py, pinned by finite differences in test_mlp_oracle.py).

Tolerances: the matrix-core product sums BF16 products in FP32, so against a
float64 product of the same BF16 values it agrees to ~K * 2^-24 relative (1e-5
here). The forward pass and the gradient requantise FP32 intermediates to BF16;
a different FP32 summation order can move a value across a BF16 rounding
boundary, changing it by one BF16 ulp (2^-8 relative) at a rate of ~1e-4 of the
elements, so whole-layer comparisons use relative norms of 5e-3."""
import numpy as np
import pytest

from oracle import mlp_oracle as mo


def _gemm(M, N, K, ta, tb, seed):
    from cmf import _native
    lib = _native.hip()
    rng = np.random.default_rng(seed)
    a = rng.normal(size=(K, M) if ta else (M, K)).astype(np.float32)
    b = rng.normal(size=(N, K) if tb else (K, N)).astype(np.float32)
    abits, bbits = mo.bf16_bits(a), mo.bf16_bits(b)
    c = np.zeros((M, N), np.float32)
    rc = lib.cmf_mlp_gemm_test(M, N, K, abits.ctypes.data, a.shape[1], int(ta), bbits.ctypes.data, b.shape[1],
                               int(tb), c.ctypes.data)
    assert rc == 0
    A = mo.bf16(a).astype(np.float64)
    B = mo.bf16(b).astype(np.float64)
    A = A.T if ta else A
    B = B.T if tb else B
    return c, A @ B


@pytest.mark.gpu
@pytest.mark.parametrize("shape", [(64, 64, 32), (70, 50, 40), (256, 2, 256), (24, 256, 1000), (1, 256, 24)])
@pytest.mark.parametrize("ta,tb", [(0, 0), (1, 0), (0, 1), (1, 1)])
def test_gemm_matches_float64(shape, ta, tb):
    M, N, K = shape
    print(f"gemm {shape} ta={ta} tb={tb}", flush=True)
    c, ref = _gemm(M, N, K, ta, tb, seed=M + N + K)
    err = np.abs(c - ref).max() / np.abs(ref).max()
    if err > 1e-5:
        # layout diagnostics: which permutation of the output tile would match
        print("max rel err", err, "| vs transposed 16x16 blocks:",
              np.abs(c - ref).max(), flush=True)
    assert err <= 1e-5


def _net(seed, rows=300, n_in=24, hid=256, n_out=2):
    rng = np.random.default_rng(seed)
    p = mo.init_params(n_in, hid, n_out, seed)
    p[-n_out:] = rng.normal(0, 0.1, n_out)
    x = rng.normal(size=(rows, n_in)).astype(np.float32)
    t = rng.normal(size=(rows, n_out)).astype(np.float32)
    return p, x, t


@pytest.mark.gpu
def test_forward_matches_oracle():
    from cmf.surrogate import MLP
    p, x, _ = _net(1)
    net = MLP(max_batch=128)            # several internal batches
    net.set_params(p)
    y = net.forward(x)
    yo, _ = mo.forward(p, x, 24, 256, 2)
    err = np.linalg.norm(y - yo) / np.linalg.norm(yo)
    print(f"forward rel err {err:.2e}", flush=True)
    assert err < 5e-3
    net.close()


@pytest.mark.gpu
@pytest.mark.parametrize("w_over", [1.0, 4.0])
def test_gradient_matches_oracle(w_over):
    from cmf.surrogate import MLP
    p, x, t = _net(2)
    net = MLP(max_batch=512)
    net.set_params(p)
    net.set_loss(w_over)
    loss = net.grad(x, t)
    g = net.grads()
    lo, go = mo.loss_grad(p, x, t, 24, 256, 2, w_over=w_over)
    assert abs(loss - lo) <= 1e-4 * abs(lo)
    o = 0
    for l, (di, do) in enumerate(mo.layer_dims(24, 256, 2)):
        for name, size in (("W", di * do), ("b", do)):
            a, b = g[o:o + size], go[o:o + size]
            err = np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30)
            print(f"layer {l} d{name}: rel err {err:.2e}", flush=True)
            assert err < 5e-3, f"layer {l} d{name}"
            o += size
    net.close()


@pytest.mark.gpu
def test_training_fits_smooth_function():
    from cmf.surrogate import MLP
    rng = np.random.default_rng(4)
    x = rng.uniform(-1, 1, size=(8192, 24)).astype(np.float32)
    t = np.stack([np.sin(2 * x[:, 0]) + x[:, 1] * x[:, 2], np.cos(x[:, 3]) * x[:, 4]], axis=1).astype(np.float32)
    net = MLP(max_batch=1024, seed=3)
    hist = net.fit(x, t, epochs=30, batch=256, lr=2e-3, seed=0)
    print(f"loss {hist[0]:.4f} -> {hist[-1]:.5f}", flush=True)
    assert hist[-1] < 0.05 * hist[0]
    net.close()
