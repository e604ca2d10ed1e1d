"""Pin the MLP oracle itself: FP64 gradients against central differences."""
import numpy as np

import mlp_ref as R


def test_reference_gradients_match_finite_differences():
    rng = np.random.default_rng(0)
    dims = [16, 32, 32, 16]
    Ws = [rng.standard_normal((dims[i], dims[i + 1])) * 0.3 for i in range(3)]
    bs = [rng.standard_normal(dims[i + 1]) * 0.1 for i in range(3)]
    X = rng.standard_normal((48, 16))
    tgt = np.stack([rng.standard_normal(48), (rng.uniform(size=48) > 0.5).astype(float)], 1)
    mask = np.ones((48, 2))

    def loss_of(Ws, bs):
        out, acts, zs = R.forward(X, Ws, bs, quantize=False)
        return R.h0_loss(out, tgt, mask)[0] / 48

    out, acts, zs = R.forward(X, Ws, bs, quantize=False)
    _, dout = R.h0_loss(out, tgt, mask)
    gWs, gbs = R.backward(dout, Ws, acts, zs, quantize=False)
    h = 1e-6
    for li in range(3):
        for (i, j) in [(0, 0), (3, 5), (7, 1)]:
            Wp = [w.copy() for w in Ws]
            Wm = [w.copy() for w in Ws]
            Wp[li][i, j] += h
            Wm[li][i, j] -= h
            fd = (loss_of(Wp, bs) - loss_of(Wm, bs)) / (2 * h)
            assert abs(fd - gWs[li][i, j]) <= 1e-6 * (1 + abs(fd)), (li, i, j)
        bp = [b.copy() for b in bs]
        bm = [b.copy() for b in bs]
        bp[li][2] += h
        bm[li][2] -= h
        fd = (loss_of(Ws, bp) - loss_of(Ws, bm)) / (2 * h)
        assert abs(fd - gbs[li][2]) <= 1e-6 * (1 + abs(fd))


def test_bf16_rounding_is_nearest_even():
    x = np.array([1.0, 1.00390625, 1.005859375, -3.0e38, 1e-40], dtype=np.float32)
    y = R.bf16(x)
    assert y[0] == 1.0
    assert y[1] == 1.0            # exactly half an ulp above 1.0: ties to even
    assert y[2] == 1.0078125      # above half: rounds up
    assert np.isfinite(y[3])


def test_mse_heads_gradient_is_the_derivative():
    rng = np.random.default_rng(3)
    M, live, width, heads = 5, 3, 4, 2
    out = rng.standard_normal((M, heads * width + 2))
    tgt = rng.standard_normal((M, live))
    mask = (rng.random((M, heads)) < 0.6).astype(float)
    s, d = R.mse_heads(out, tgt, heads, width, mask)
    h = 1e-6
    for i in range(M):
        for j in range(out.shape[1]):
            o = out.copy()
            o[i, j] += h
            fd = (R.mse_heads(o, tgt, heads, width, mask)[0] - s) / h / (M * live)
            assert abs(fd - d[i, j]) <= 1e-5 * (1 + abs(fd)), (i, j)
    # padding columns and masked-out rows carry no gradient
    assert np.all(d[:, live:width] == 0) and np.all(d[:, heads * width:] == 0)
    for i in range(M):
        for hh in range(heads):
            if mask[i, hh] == 0:
                assert np.all(d[i, hh * width:hh * width + live] == 0)
