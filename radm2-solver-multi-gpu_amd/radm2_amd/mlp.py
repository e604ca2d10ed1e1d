"""Small MLPs on the project's own BF16-MFMA kernels (include/radm2_mlp.h).

A network is a stack of dense layers; every layer's forward, backward and
Adam update is one of the HIP kernels in csrc/radm2_mlp.hip.  torch is used
only to own device buffers and for host-side bookkeeping (slicing, means of
features); no torch arithmetic touches the forward or backward pass.

Shapes: inputs are padded to a multiple of 16 features, outputs to 16
columns; the batch is a multiple of 16.
"""
import ctypes

import numpy as np

from . import load_library

ACT_NONE = 0
ACT_SILU = 1
NFEAT = 83
FPAD = 96

_P = ctypes.c_void_p
_SIGS = {
    "radm2_mlp_dense_fwd": [_P, ctypes.c_int, ctypes.c_int, ctypes.c_int, _P, _P, _P, ctypes.c_int,
                            _P, _P, _P],
    "radm2_mlp_act_bwd": [_P, ctypes.c_int64, _P, _P, ctypes.c_int, _P, _P],
    "radm2_mlp_dense_bwd": [_P, ctypes.c_int, ctypes.c_int, ctypes.c_int, _P, _P, _P, _P, _P, _P, _P],
    "radm2_mlp_h0_loss": [_P, ctypes.c_int, ctypes.c_int, _P, _P, _P, _P, _P],
    "radm2_mlp_mse_loss": [_P, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                           _P, _P, _P, _P, _P],
    "radm2_mlp_adam": [_P, ctypes.c_int64, _P, _P, _P, _P, _P, ctypes.c_float, ctypes.c_float,
                       ctypes.c_float, ctypes.c_float, ctypes.c_int, ctypes.c_float],
    "radm2_mlp_to_bf16": [_P, ctypes.c_int64, _P, _P],
    "radm2_mlp_features": [_P, ctypes.c_int64, ctypes.c_int64, _P, _P, _P, _P, _P, _P, _P,
                           ctypes.c_int, _P, _P],
    "radm2_mlp_h0_select": [_P, ctypes.c_int64, ctypes.c_int, _P, ctypes.c_double, ctypes.c_double,
                            ctypes.c_float, ctypes.c_double, _P, _P],
    "radm2_mlp_dlog_targets": [_P, ctypes.c_int64, ctypes.c_int64, _P, _P, ctypes.c_double, _P, _P,
                               _P],
    "radm2_mlp_surrogate_apply": [_P, ctypes.c_int64, ctypes.c_int64, ctypes.c_int, ctypes.c_int, _P,
                                  _P, _P, ctypes.c_double, ctypes.c_double, ctypes.c_float,
                                  ctypes.c_double, ctypes.c_uint64, _P, _P, _P],
    "radm2_mlp_compact": [_P, ctypes.c_int64, _P, _P, _P],
    "radm2_mlp_gather": [_P, ctypes.c_int, ctypes.c_int, ctypes.c_int64, _P, ctypes.c_int64, _P,
                         ctypes.c_int64, _P],
    "radm2_mlp_scatter": [_P, ctypes.c_int, ctypes.c_int, ctypes.c_int64, _P, ctypes.c_int64, _P,
                          ctypes.c_int64, _P],
}
EXPORTED_SYMBOLS = tuple(_SIGS)
_lib = None


def lib():
    global _lib
    if _lib is None:
        L = load_library()
        for name, args in _SIGS.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = ctypes.c_int
        _lib = L
    return _lib


def _p(t):
    return None if t is None else ctypes.c_void_p(t.data_ptr())


def _st():
    """torch's current stream: every kernel here is ordered with torch's own
    operations on the same buffers."""
    import torch
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _ck(code):
    if code != 0:
        from . import Radm2Error, strerror
        raise Radm2Error(code, strerror(code))


def bf16_bits(x):
    """numpy float32 -> bf16 bit pattern (uint16), round to nearest even."""
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return u.astype(np.uint16)


def bf16_round(x):
    """numpy float32 rounded through bf16 (what the kernels see)."""
    return (bf16_bits(x).astype(np.uint32) << 16).view(np.float32)


class Dense:
    """One layer: W [K x N] (FP32 master + BF16 copy), b [N], activation."""

    def __init__(self, K, N, act, rng, device="cuda"):
        import torch
        self.K, self.N, self.act = K, N, act
        w = (rng.standard_normal((K, N)) * np.sqrt(2.0 / K)).astype(np.float32)
        self.w = torch.from_numpy(w).to(device)
        self.wb = torch.empty((K, N), dtype=torch.int16, device=device)
        self.b = torch.zeros(N, dtype=torch.float32, device=device)
        self.gw = torch.zeros_like(self.w)
        self.gb = torch.zeros_like(self.b)
        self.mw, self.vw = torch.zeros_like(self.w), torch.zeros_like(self.w)
        self.mb, self.vb = torch.zeros_like(self.b), torch.zeros_like(self.b)
        self.sync_bf16()

    def sync_bf16(self):
        _ck(lib().radm2_mlp_to_bf16(_st(), self.w.numel(), _p(self.w), _p(self.wb)))

    def load(self, w, b):
        import torch
        self.w.copy_(torch.from_numpy(np.ascontiguousarray(w, dtype=np.float32)))
        self.b.copy_(torch.from_numpy(np.ascontiguousarray(b, dtype=np.float32)))
        self.sync_bf16()


class MLP:
    """Dense stack on the project's kernels.  dims: [in, h1, ..., out] with
    every entry a multiple of 16; hidden layers SiLU, last layer linear."""

    def __init__(self, dims, seed=0, device="cuda"):
        rng = np.random.default_rng(seed)
        assert all(d % 16 == 0 for d in dims)
        self.dims = list(dims)
        self.layers = [Dense(dims[i], dims[i + 1], ACT_SILU if i + 2 < len(dims) else ACT_NONE,
                             rng, device) for i in range(len(dims) - 1)]
        self.step = 0
        self.device = device
        self._cache = None

    # -- forward / backward -------------------------------------------------
    def forward(self, Xb, keep=False):
        """Xb: [M x dims[0]] int16 (BF16 bits) device tensor; returns the
        FP32 output [M x dims[-1]]."""
        import torch
        M = Xb.shape[0]
        acts = [Xb]
        zs = []
        x = Xb
        out = None
        for i, L in enumerate(self.layers):
            last = i == len(self.layers) - 1
            Z = torch.empty((M, L.N), dtype=torch.float32, device=self.device) if keep else None
            if last:
                out = torch.empty((M, L.N), dtype=torch.float32, device=self.device)
                _ck(lib().radm2_mlp_dense_fwd(_st(), M, L.K, L.N, _p(x), _p(L.wb), _p(L.b), L.act,
                                              _p(Z), None, _p(out)))
            else:
                yb = torch.empty((M, L.N), dtype=torch.int16, device=self.device)
                _ck(lib().radm2_mlp_dense_fwd(_st(), M, L.K, L.N, _p(x), _p(L.wb), _p(L.b), L.act,
                                              _p(Z), _p(yb), None))
                x = yb
                acts.append(yb)
            zs.append(Z)
        if keep:
            self._cache = (acts, zs)
        return out

    def backward(self, dout):
        """Accumulate weight gradients from d(loss)/d(output) (FP32)."""
        import torch
        acts, zs = self._cache
        M = dout.shape[0]
        dY = dout
        for i in range(len(self.layers) - 1, -1, -1):
            L = self.layers[i]
            dZb = torch.empty((M, L.N), dtype=torch.int16, device=self.device)
            dZf = torch.empty((M, L.N), dtype=torch.float32, device=self.device)
            _ck(lib().radm2_mlp_act_bwd(_st(), M * L.N, _p(dY), _p(zs[i]), L.act, _p(dZb), _p(dZf)))
            dX = torch.empty((M, L.K), dtype=torch.float32, device=self.device) if i > 0 else None
            _ck(lib().radm2_mlp_dense_bwd(_st(), M, L.K, L.N, _p(acts[i]), _p(L.wb), _p(dZb),
                                          _p(dZf), _p(dX), _p(L.gw), _p(L.gb)))
            dY = dX

    def zero_grad(self):
        for L in self.layers:
            L.gw.zero_()
            L.gb.zero_()

    def adam(self, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8, wd=0.0):
        self.step += 1
        for L in self.layers:
            _ck(lib().radm2_mlp_adam(_st(), L.w.numel(), _p(L.w), _p(L.wb), _p(L.gw), _p(L.mw),
                                     _p(L.vw), lr, b1, b2, eps, self.step, wd))
            _ck(lib().radm2_mlp_adam(_st(), L.b.numel(), _p(L.b), None, _p(L.gb), _p(L.mb),
                                     _p(L.vb), lr, b1, b2, eps, self.step, 0.0))

    # -- persistence ----------------------------------------------------------
    def state(self):
        d = {"dims": np.array(self.dims)}
        for i, L in enumerate(self.layers):
            d["w%d" % i] = L.w.cpu().numpy()
            d["b%d" % i] = L.b.cpu().numpy()
        return d

    def load_state(self, d):
        assert list(d["dims"]) == self.dims
        for i, L in enumerate(self.layers):
            L.load(d["w%d" % i], d["b%d" % i])


def features(batch, mu, sd, ncells=None, F=FPAD, want_f32=False):
    """Standardised BF16 features [n x F] of a CellBatch's current state."""
    import torch
    n = batch.ncells if ncells is None else ncells
    X = torch.empty((n, F), dtype=torch.int16, device=batch.conc.device)
    Xf = torch.empty((n, F), dtype=torch.float32, device=batch.conc.device) if want_f32 else None
    _ck(lib().radm2_mlp_features(_st(), n, batch.ld, _p(batch.conc), _p(batch.temp), _p(batch.cair),
                                 _p(batch.h2o), _p(batch.jrate), _p(mu), _p(sd), F, _p(X), _p(Xf)))
    return X, Xf
