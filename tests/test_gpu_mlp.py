"""BF16-MFMA MLP kernels against the CPU oracle (tests/mlp_ref.py)."""
import numpy as np
import pytest

import mlp_ref as R

pytestmark = pytest.mark.gpu


def _t(a, dtype=None):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def test_mfma_dense_forward_layout():
    import torch
    from radm2_amd import mlp
    rng = np.random.default_rng(1)
    M, K, N = 48, 32, 48
    X = rng.standard_normal((M, K)).astype(np.float32)
    W = rng.standard_normal((K, N)).astype(np.float32)
    b = rng.standard_normal(N).astype(np.float32)
    Xb = _t(mlp.bf16_bits(X).view(np.int16))
    Wb = _t(mlp.bf16_bits(W).view(np.int16))
    out = torch.empty((M, N), dtype=torch.float32, device="cuda")
    bt = _t(b)  # keep every operand alive until the kernel has run
    mlp._ck(mlp.lib().radm2_mlp_dense_fwd(mlp._st(), M, K, N, mlp._p(Xb), mlp._p(Wb), mlp._p(bt),
                                          mlp.ACT_NONE, None, None, mlp._p(out)))
    torch.cuda.synchronize()
    ref = R.bf16(X) @ R.bf16(W) + b
    np.testing.assert_allclose(out.cpu().numpy(), ref, rtol=1e-5, atol=1e-5)


def test_forward_backward_match_oracle():
    import torch
    from radm2_amd import mlp
    rng = np.random.default_rng(2)
    dims = [96, 128, 128, 16]
    net = mlp.MLP(dims, seed=3)
    for L in net.layers:
        L.b.copy_(_t(rng.standard_normal(L.N).astype(np.float32) * 0.1))
    M = 1024 + 64
    X = rng.standard_normal((M, 96)).astype(np.float32)
    Xb = _t(mlp.bf16_bits(X).view(np.int16))
    out = net.forward(Xb, keep=True)
    Ws = [L.w.cpu().numpy().astype(np.float64) for L in net.layers]
    bs = [L.b.cpu().numpy().astype(np.float64) for L in net.layers]
    rout, acts, zs = R.forward(X, Ws, bs)
    np.testing.assert_allclose(out.cpu().numpy(), rout, rtol=2e-3, atol=2e-3)
    tgt = np.stack([rng.standard_normal(M), (rng.uniform(size=M) > 0.5)], 1).astype(np.float32)
    mask = np.ones((M, 2), np.float32)
    dout = torch.zeros((M, 16), dtype=torch.float32, device="cuda")
    loss = torch.zeros(1, dtype=torch.float32, device="cuda")
    tt, tm = _t(tgt), _t(mask)  # keep alive: a dropped temporary's block is reused at once
    mlp._ck(mlp.lib().radm2_mlp_h0_loss(mlp._st(), M, 16, mlp._p(out), mlp._p(tt), mlp._p(tm),
                                        mlp._p(dout), mlp._p(loss)))
    torch.cuda.synchronize()
    rl, rd = R.h0_loss(out.cpu().numpy().astype(np.float64), tgt, mask)
    assert loss.item() == pytest.approx(rl, rel=1e-4)
    np.testing.assert_allclose(dout.cpu().numpy(), rd, rtol=1e-4, atol=1e-9)
    net.zero_grad()
    net.backward(dout)
    # oracle backward from the GPU's own forward state, bf16 at the same points
    gWs, gbs = R.backward(dout.cpu().numpy().astype(np.float64), Ws, acts, zs)
    for L, gw, gb in zip(net.layers, gWs, gbs):
        g = L.gw.cpu().numpy()
        scale = np.max(np.abs(gw))
        assert np.max(np.abs(g - gw)) <= 2e-2 * scale
        assert np.corrcoef(g.ravel(), gw.ravel())[0, 1] > 0.999
        np.testing.assert_allclose(L.gb.cpu().numpy(), gb, rtol=2e-2, atol=2e-2 * np.max(np.abs(gb)))


def test_adam_matches_numpy():
    import torch
    from radm2_amd import mlp
    rng = np.random.default_rng(4)
    w = rng.standard_normal(1000).astype(np.float32)
    g = rng.standard_normal(1000).astype(np.float32)
    tw, tg = _t(w.copy()), _t(g)
    m1 = torch.zeros(1000, dtype=torch.float32, device="cuda")
    m2 = torch.zeros_like(m1)
    mlp._ck(mlp.lib().radm2_mlp_adam(mlp._st(), 1000, mlp._p(tw), None, mlp._p(tg), mlp._p(m1),
                                     mlp._p(m2), 1e-2, 0.9, 0.999, 1e-8, 1, 0.0))
    mh = 0.1 * g / 0.1
    vh = 0.001 * g * g / 0.001
    ref = w - 1e-2 * mh / (np.sqrt(vh) + 1e-8)
    np.testing.assert_allclose(tw.cpu().numpy(), ref, rtol=1e-5, atol=1e-6)


def test_features_match_numpy():
    import torch
    import radm2_amd
    from radm2_amd import mlp, synth
    c = synth.make_cells(300, 9, 100.0, 0.05)
    d = synth.to_device(c)
    b = radm2_amd.CellBatch(d["conc"], d["temp"], d["cair"], d["h2o"], d["jrate"], rc_n2o5=d["rc_n2o5"])
    mu = torch.zeros(83, dtype=torch.float32, device="cuda")
    sd = torch.ones(83, dtype=torch.float32, device="cuda")
    X, Xf = mlp.features(b, mu, sd, want_f32=True)
    f = np.concatenate([np.log10(np.maximum(c["conc"], 1e-30)).T, (c["temp"] / 300.0)[:, None],
                        (np.log10(c["cair"]) - 19.0)[:, None], (np.log10(c["h2o"]) - 17.0)[:, None],
                        (np.log10(c["jrate"] + 1e-12) + 6.0).T], axis=1)
    got = Xf.cpu().numpy()
    np.testing.assert_allclose(got[:, :83], f, rtol=1e-5, atol=1e-5)
    assert (got[:, 83:] == 0).all()
    np.testing.assert_array_equal(X.cpu().numpy().view(np.uint16), mlp.bf16_bits(got))


def test_training_reduces_loss():
    import torch
    from radm2_amd import mlp
    rng = np.random.default_rng(5)
    M = 4096
    X = rng.standard_normal((M, 96)).astype(np.float32)
    X[:, 83:] = 0
    y = np.tanh(X[:, 0] - 0.5 * X[:, 1]) * 2 + 0.3 * X[:, 2]
    tgt = np.stack([y, (y > 0)], 1).astype(np.float32)
    mask = np.ones((M, 2), np.float32)
    net = mlp.MLP([96, 128, 128, 16], seed=6)
    Xb = _t(mlp.bf16_bits(X).view(np.int16))
    tt, tm = _t(tgt), _t(mask)
    losses = []
    for it in range(150):
        out = net.forward(Xb, keep=True)
        dout = torch.zeros_like(out)
        loss = torch.zeros(1, dtype=torch.float32, device="cuda")
        mlp._ck(mlp.lib().radm2_mlp_h0_loss(mlp._st(), M, 16, mlp._p(out), mlp._p(tt), mlp._p(tm),
                                            mlp._p(dout), mlp._p(loss)))
        net.zero_grad()
        net.backward(dout)
        net.adam(lr=3e-3)
        losses.append(loss.item() / M)
    assert losses[-1] < 0.2 * losses[0], (losses[0], losses[-1])
