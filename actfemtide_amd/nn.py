"""Learned warm start: a 3-D convolutional encoder-decoder surrogate, trained and run on the
library's own BF16-MFMA kernels (include/afx_nn.h).

The reference has no learned component (no neural, surrogate or learned code anywhere in its
tree). The surrogate only supplies the initial guess X that COCG13_single takes on entry
(src/solver/COCG13_single.f90:9-36, X is in/out): the Krylov solve always finishes to the
same tolerance, so a poor prediction costs iterations, never accuracy.

Network (4 resolution levels, channels c, 2c, 4c, 8c, ReLU after every layer but the head):
    e0a: conv3 in->c, e0b: conv3 c->c                                   level 0 (skip 0)
    d0: conv 2x2x2/2 c->2c, e1: conv3 2c->2c                            level 1 (skip 1)
    d1: conv 2x2x2/2 2c->4c, e2: conv3 4c->4c                           level 2 (skip 2)
    d2: conv 2x2x2/2 4c->8c, b3: conv3 8c->8c                           level 3
    u2: transposed 2x2x2/2 8c->4c, + skip 2, c2: conv3 4c->4c
    u1: transposed 2x2x2/2 4c->2c, + skip 1, c1: conv3 2c->2c
    u0: transposed 2x2x2/2 2c->c,  + skip 0, c0: conv3 c->c
    head: 1x1 conv c->6 (Re/Im of a cell vector field), FP32 out
Weights are FP32 masters with BF16 copies; activations BF16 channels-last; accumulation FP32;
Adam on the masters. Every operation is a libafx kernel: no PyTorch operator computes any of
it (torch tensors only hold device memory).

Cell field <-> edge values: the network predicts, per cell, a constant complex vector F; the
edge value is x_e = F . t_e with t_e the edge vector (line(1,e) -> line(2,e)), averaged over the
cells that contain the edge (afx_nn_cells_to_edges). Training targets are the least-squares
constant fields of solved edge values per cell (afx_nn_edges_to_cells).
"""
from __future__ import annotations

import ctypes as C
import math
import os

import numpy as np

from . import afx

_VP, _I, _D, _F, _LL = C.c_void_p, C.c_int, C.c_double, C.c_float, C.c_longlong
SIGNATURES = {
    "afx_nn_last_error": ([], C.c_char_p),
    "afx_nn_gemm_nt": ([_VP, _I, _VP, _I, _I, _I, _I, _VP, _I, _VP, _I, _VP, _I, _I, _I, _VP, _VP], _I),
    "afx_nn_im2col3": ([_VP, _VP, _I, _I, _I, _I, _I, _I, _I, _VP], _I),
    "afx_nn_im2col3_t": ([_VP, _VP, _I, _I, _I, _I, _I, _I, _I, _VP], _I),
    "afx_nn_weight_t3": ([_VP, _I, _VP, _I, _I, _I, _VP], _I),
    "afx_nn_space_to_depth": ([_VP, _VP, _I, _I, _I, _I, _I, _I, _I, _VP], _I),
    "afx_nn_transpose": ([_VP, _I, _I, _I, _I, _VP, _I, _VP], _I),
    "afx_nn_cast_f2b": ([_VP, _VP, _LL, _VP], _I),
    "afx_nn_relu_bwd": ([_VP, _VP, _VP, _VP, _LL, _VP], _I),
    "afx_nn_colsum": ([_VP, _I, _I, _VP, _I, _VP, _VP], _I),
    "afx_nn_add_bf16": ([_VP, _VP, _VP, _LL, _VP], _I),
    "afx_nn_add_f32": ([_VP, _VP, _VP, _LL, _VP], _I),
    "afx_nn_mse": ([_VP, _I, _VP, _VP, _LL, _I, _F, _VP, _VP, _VP, _VP], _I),
    "afx_nn_adam": ([_VP, _VP, _VP, _VP, _VP, _LL, _F, _F, _F, _F, _I, _VP], _I),
    "afx_nn_cells_to_edges": ([_VP, _VP, _VP, _VP, _I, _D, _VP, _I, _VP, _VP], _I),
    "afx_nn_edges_to_cells": ([_VP, _VP, _VP, _VP, _D, _I, _VP, _I, _VP], _I),
}

_nn = None


def lib():
    """libafx.so with the afx_nn_* prototypes set (raises when the library is missing)."""
    global _nn
    if _nn is None:
        L = afx.lib()
        for name, (args, res) in SIGNATURES.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = res
        _nn = L
    return _nn


class NNError(RuntimeError):
    pass


def _call(name, *args):
    rc = getattr(lib(), name)(*args)
    if rc != 0:
        raise NNError(f"{name}: {rc} {lib().afx_nn_last_error().decode()}")


def _p(t):
    return None if t is None else C.c_void_p(t.data_ptr())


def _ru(x, m):
    return (x + m - 1) // m * m


def _torch():
    import torch
    return torch


def _stream():
    return C.c_void_p(_torch().cuda.current_stream().cuda_stream)


# ----------------------------------------------------------------------------------------------
# primitive wrappers (device tensors in, device tensors out)

def gemm_nt(A, B, M, N, K, C_out=None, Cb=None, bias=None, relu=False, splits=None, lda=None, ldb=None, ldc=None,
            ldcb=None, bias_period=0):
    """C[m][n] = sum_k A[m][k] B[n][k] (+bias) (ReLU) on the BF16 MFMA GEMM. splits=None splits K
    when the output has fewer 128-wide tiles than the GPU has CUs and K is long (the deep levels),
    so that every CU has work; the partial sums are added in split order (deterministic)."""
    torch = _torch()
    lda = lda or A.shape[-1]
    ldb = ldb or B.shape[-1]
    if splits is None:
        nw = 4 if N > 64 else (2 if N > 32 else 1)
        tiles = ((M + 127) // 128) * ((N + 32 * nw - 1) // (32 * nw))
        splits = max(1, min(K // 256, -(-1024 // tiles))) if (tiles < 256 and K >= 1024) else 1
    part = None
    if splits > 1:
        part = torch.empty(splits * M * N, dtype=torch.float32, device=A.device)
    _call("afx_nn_gemm_nt", _p(A), lda, _p(B), ldb, M, N, K, _p(C_out), ldc or (C_out.shape[-1] if C_out is not None else 0),
          _p(Cb), ldcb or (Cb.shape[-1] if Cb is not None else 0), _p(bias), bias_period, 1 if relu else 0, splits,
          _p(part), _stream())


def colsum(X, R, Cc, out, accumulate=False):
    """out[c] (+)= sum_r X[r][c] for an f32 [R][Cc] device array (two deterministic passes)."""
    torch = _torch()
    part = torch.empty(((R + 255) // 256) * Cc, dtype=torch.float32, device=out.device)
    _call("afx_nn_colsum", _p(X), R, Cc, _p(out), 1 if accumulate else 0, _p(part), _stream())


def _splits(M, N, K):
    tiles = ((M + 127) // 128) * ((N + 127) // 128)
    s = max(1, min(512 // max(tiles, 1), K // 512))
    return s


def transpose(X, R, Cc, Rp, ldx=None):
    """bf16 [Cc][Rp] transpose of X [R][ldx] (f32 or bf16), zero padded."""
    torch = _torch()
    Y = torch.empty((Cc, Rp), dtype=torch.bfloat16, device=X.device)
    es = 4 if X.dtype == torch.float32 else 2
    _call("afx_nn_transpose", _p(X), es, ldx or X.shape[-1], R, Cc, _p(Y), Rp, _stream())
    return Y


def cast_bf16(x):
    torch = _torch()
    y = torch.empty(x.shape, dtype=torch.bfloat16, device=x.device)
    _call("afx_nn_cast_f2b", _p(x), _p(y), x.numel(), _stream())
    return y


# ----------------------------------------------------------------------------------------------
# layers

class _Param:
    def __init__(self, w: np.ndarray, device):
        torch = _torch()
        self.w = torch.from_numpy(np.ascontiguousarray(w, np.float32)).to(device)
        self.g = torch.zeros_like(self.w)
        self.m = torch.zeros_like(self.w)
        self.v = torch.zeros_like(self.w)
        self.wb = cast_bf16(self.w) if device.type == "cuda" else None

    def refresh(self):
        self.wb = cast_bf16(self.w)


class Conv3:
    """3x3x3 convolution, padding 1, ci -> co, ReLU. W [co][27*ci padded to kp]."""

    def __init__(self, ci, co, rng, device):
        self.ci, self.co = ci, co
        self.kp = _ru(27 * ci, 32)
        self.kpt = _ru(27 * co, 32)
        w = np.zeros((co, self.kp), np.float32)
        w[:, :27 * ci] = rng.standard_normal((co, 27 * ci)) * math.sqrt(2.0 / (27 * ci))
        self.W = _Param(w, device)
        self.b = _Param(np.zeros(co, np.float32), device)

    def params(self):
        return [self.W, self.b]

    def forward(self, x, dims):
        torch = _torch()
        nb, D, H, W = dims
        V = nb * D * H * W
        cols = torch.empty((V, self.kp), dtype=torch.bfloat16, device=x.device)
        _call("afx_nn_im2col3", _p(x), _p(cols), nb, D, H, W, self.ci, self.kp, 0, _stream())
        y = torch.empty((V, self.co), dtype=torch.bfloat16, device=x.device)
        gemm_nt(cols, self.W.wb, V, self.co, self.kp, Cb=y, bias=self.b.w, relu=True)
        self.saved = (x, y, dims)
        return y

    def backward(self, dy, need_dx=True):
        torch = _torch()
        x, y, dims = self.saved
        nb, D, H, W = dims
        V = nb * D * H * W
        co, ci = self.co, self.ci
        dP = torch.empty((V, co), dtype=torch.float32, device=dy.device)
        dPb = torch.empty((V, co), dtype=torch.bfloat16, device=dy.device)
        _call("afx_nn_relu_bwd", _p(dy), _p(y), _p(dP), _p(dPb), V * co, _stream())
        colsum(dP, V, co, self.b.g)
        Vp = _ru(V, 32)
        dPt = transpose(dP, V, co, Vp)
        colsT = torch.empty((self.kp, Vp), dtype=torch.bfloat16, device=dy.device)
        _call("afx_nn_im2col3_t", _p(x), _p(colsT), nb, D, H, W, ci, self.kp, Vp, _stream())
        gemm_nt(dPt, colsT, co, self.kp, Vp, C_out=self.W.g, splits=_splits(co, self.kp, Vp))
        if not need_dx:
            return None
        colsF = torch.empty((V, self.kpt), dtype=torch.bfloat16, device=dy.device)
        _call("afx_nn_im2col3", _p(dPb), _p(colsF), nb, D, H, W, co, self.kpt, 1, _stream())
        wt = torch.empty((ci, self.kpt), dtype=torch.bfloat16, device=dy.device)
        _call("afx_nn_weight_t3", _p(self.W.wb), self.kp, _p(wt), co, ci, self.kpt, _stream())
        dx = torch.empty((V, ci), dtype=torch.float32, device=dy.device)
        gemm_nt(colsF, wt, V, ci, self.kpt, C_out=dx)
        return dx


class Down:
    """2x2x2 convolution with stride 2, ci -> co, ReLU. W [co][8*ci] (sub-voxel major)."""

    def __init__(self, ci, co, rng, device):
        self.ci, self.co = ci, co
        self.W = _Param(rng.standard_normal((co, 8 * ci)) * math.sqrt(2.0 / (8 * ci)), device)
        self.b = _Param(np.zeros(co, np.float32), device)

    def params(self):
        return [self.W, self.b]

    def forward(self, x, dims):
        torch = _torch()
        nb, D, H, W = dims
        Vo = nb * D * H * W // 8
        S = torch.empty((Vo, 8 * self.ci), dtype=torch.bfloat16, device=x.device)
        _call("afx_nn_space_to_depth", _p(x), _p(S), nb, D, H, W, self.ci, 2, 0, _stream())
        y = torch.empty((Vo, self.co), dtype=torch.bfloat16, device=x.device)
        gemm_nt(S, self.W.wb, Vo, self.co, 8 * self.ci, Cb=y, bias=self.b.w, relu=True)
        self.saved = (S, y, dims)
        return y, (nb, D // 2, H // 2, W // 2)

    def backward(self, dy):
        torch = _torch()
        S, y, dims = self.saved
        nb, D, H, W = dims
        Vo = nb * D * H * W // 8
        co, ci = self.co, self.ci
        dP = torch.empty((Vo, co), dtype=torch.float32, device=dy.device)
        dPb = torch.empty((Vo, co), dtype=torch.bfloat16, device=dy.device)
        _call("afx_nn_relu_bwd", _p(dy), _p(y), _p(dP), _p(dPb), Vo * co, _stream())
        colsum(dP, Vo, co, self.b.g)
        Vp = _ru(Vo, 32)
        dPt = transpose(dP, Vo, co, Vp)
        St = transpose(S, Vo, 8 * ci, Vp)
        gemm_nt(dPt, St, co, 8 * ci, Vp, C_out=self.W.g, splits=_splits(co, 8 * ci, Vp))
        wT = transpose(self.W.wb, co, 8 * ci, co)
        dS = torch.empty((Vo, 8 * ci), dtype=torch.float32, device=dy.device)
        gemm_nt(dPb, wT, Vo, 8 * ci, co, C_out=dS)
        dx = torch.empty((nb * D * H * W, ci), dtype=torch.float32, device=dy.device)
        _call("afx_nn_space_to_depth", _p(dx), _p(dS), nb, D, H, W, ci, 4, 1, _stream())
        return dx


class Up:
    """Transposed 2x2x2 convolution with stride 2, ci -> co, ReLU. W [8*co][ci]; the bias is per
    output channel (shared by the 8 sub-voxels)."""

    def __init__(self, ci, co, rng, device):
        self.ci, self.co = ci, co
        self.W = _Param(rng.standard_normal((8 * co, ci)) * math.sqrt(2.0 / ci), device)
        self.b = _Param(np.zeros(co, np.float32), device)

    def params(self):
        return [self.W, self.b]

    def forward(self, x, dims):
        torch = _torch()
        nb, D, H, W = dims
        V = nb * D * H * W
        yc = torch.empty((V, 8 * self.co), dtype=torch.bfloat16, device=x.device)
        gemm_nt(x, self.W.wb, V, 8 * self.co, self.ci, Cb=yc, bias=self.b.w, bias_period=self.co, relu=True)
        fine = (nb, 2 * D, 2 * H, 2 * W)
        y = torch.empty((8 * V, self.co), dtype=torch.bfloat16, device=x.device)
        _call("afx_nn_space_to_depth", _p(y), _p(yc), nb, 2 * D, 2 * H, 2 * W, self.co, 2, 1, _stream())
        self.saved = (x, yc, dims)
        return y, fine

    def backward(self, dy):
        torch = _torch()
        x, yc, dims = self.saved
        nb, D, H, W = dims
        V = nb * D * H * W
        co, ci = self.co, self.ci
        dyc = torch.empty((V, 8 * co), dtype=torch.float32, device=dy.device)
        _call("afx_nn_space_to_depth", _p(dy), _p(dyc), nb, 2 * D, 2 * H, 2 * W, co, 4, 0, _stream())
        dP = torch.empty((V, 8 * co), dtype=torch.float32, device=dy.device)
        dPb = torch.empty((V, 8 * co), dtype=torch.bfloat16, device=dy.device)
        _call("afx_nn_relu_bwd", _p(dyc), _p(yc), _p(dP), _p(dPb), V * 8 * co, _stream())
        colsum(dP, 8 * V, co, self.b.g)
        Vp = _ru(V, 32)
        dPt = transpose(dP, V, 8 * co, Vp)
        xt = transpose(x, V, ci, Vp)
        gemm_nt(dPt, xt, 8 * co, ci, Vp, C_out=self.W.g, splits=_splits(8 * co, ci, Vp))
        wT = transpose(self.W.wb, 8 * co, ci, 8 * co)
        dx = torch.empty((V, ci), dtype=torch.float32, device=dy.device)
        gemm_nt(dPb, wT, V, ci, 8 * co, C_out=dx)
        return dx


class Head:
    """1x1 convolution ci -> nout (linear), computed with 32 output columns (rows nout.. of W
    stay zero), FP32 output [V][32]."""

    def __init__(self, ci, rng, device, nout=6):
        self.ci, self.nout = ci, nout
        w = np.zeros((32, ci), np.float32)
        w[:nout] = rng.standard_normal((nout, ci)) * math.sqrt(1.0 / ci)
        self.W = _Param(w, device)
        self.b = _Param(np.zeros(32, np.float32), device)

    def params(self):
        return [self.W, self.b]

    def forward(self, x, dims):
        torch = _torch()
        V = x.shape[0]
        out = torch.empty((V, 32), dtype=torch.float32, device=x.device)
        gemm_nt(x, self.W.wb, V, 32, self.ci, C_out=out, bias=self.b.w)
        self.saved = (x,)
        return out

    def backward(self, dout):
        torch = _torch()
        (x,) = self.saved
        V = x.shape[0]
        colsum(dout, V, 32, self.b.g)
        Vp = _ru(V, 32)
        dt = transpose(dout, V, 32, Vp)
        xt = transpose(x, V, self.ci, Vp)
        gemm_nt(dt, xt, 32, self.ci, Vp, C_out=self.W.g, splits=_splits(32, self.ci, Vp))
        db = cast_bf16(dout)
        wT = transpose(self.W.wb, 32, self.ci, 32)
        dx = torch.empty((V, self.ci), dtype=torch.float32, device=dout.device)
        gemm_nt(db, wT, V, self.ci, 32, C_out=dx)
        return dx


def _add_b(a, b):
    torch = _torch()
    y = torch.empty_like(a)
    _call("afx_nn_add_bf16", _p(a), _p(b), _p(y), a.numel(), _stream())
    return y


def _add_f(a, b):
    torch = _torch()
    y = torch.empty_like(a)
    _call("afx_nn_add_f32", _p(a), _p(b), _p(y), a.numel(), _stream())
    return y


class Surrogate:
    """The encoder-decoder (module docstring). `base` channels at level 0 (a multiple of 16),
    `cin` input channels (a multiple of 8)."""

    NOUT = 6

    def __init__(self, base: int = 32, cin: int = 8, seed: int = 0, device="cuda"):
        torch = _torch()
        if base % 16 or cin % 8:
            raise ValueError("base must be a multiple of 16 and cin a multiple of 8")
        dev = torch.device(device)
        rng = np.random.default_rng(seed)
        c0, c1, c2, c3 = base, 2 * base, 4 * base, 8 * base
        self.base, self.cin = base, cin
        self.e0a = Conv3(cin, c0, rng, dev)
        self.e0b = Conv3(c0, c0, rng, dev)
        self.d0 = Down(c0, c1, rng, dev)
        self.e1 = Conv3(c1, c1, rng, dev)
        self.d1 = Down(c1, c2, rng, dev)
        self.e2 = Conv3(c2, c2, rng, dev)
        self.d2 = Down(c2, c3, rng, dev)
        self.b3 = Conv3(c3, c3, rng, dev)
        self.u2 = Up(c3, c2, rng, dev)
        self.c2 = Conv3(c2, c2, rng, dev)
        self.u1 = Up(c2, c1, rng, dev)
        self.c1 = Conv3(c1, c1, rng, dev)
        self.u0 = Up(c1, c0, rng, dev)
        self.c0 = Conv3(c0, c0, rng, dev)
        self.head = Head(c0, rng, dev, self.NOUT)
        self.layers = ["e0a", "e0b", "d0", "e1", "d1", "e2", "d2", "b3", "u2", "c2", "u1", "c1", "u0", "c0", "head"]
        self.step_count = 0

    def named_params(self):
        out = []
        for name in self.layers:
            L = getattr(self, name)
            out.append((name + ".W", L.W))
            out.append((name + ".b", L.b))
        return out

    def forward(self, x, dims):
        """x: bf16 [nb*D*H*W][cin] channels-last, dims = (nb, D, H, W) with D, H, W divisible
        by 8. Returns the FP32 output [V][32] (channels 0..5 used)."""
        if any(d % 8 for d in dims[1:]):
            raise ValueError("grid dimensions must be divisible by 8 (three stride-2 levels)")
        a = self.e0a.forward(x, dims)
        s0 = self.e0b.forward(a, dims)
        a, g1 = self.d0.forward(s0, dims)
        s1 = self.e1.forward(a, g1)
        a, g2 = self.d1.forward(s1, g1)
        s2 = self.e2.forward(a, g2)
        a, g3 = self.d2.forward(s2, g2)
        a = self.b3.forward(a, g3)
        a, _ = self.u2.forward(a, g3)
        a = self.c2.forward(_add_b(a, s2), g2)
        a, _ = self.u1.forward(a, g2)
        a = self.c1.forward(_add_b(a, s1), g1)
        a, _ = self.u0.forward(a, g1)
        a = self.c0.forward(_add_b(a, s0), dims)
        return self.head.forward(a, dims)

    def backward(self, dout):
        """dout: FP32 [V][32] gradient of the loss w.r.t. the output; fills every .g."""
        g = self.head.backward(dout)
        g = self.c0.backward(g)          # d(u0 out + s0)
        ds0 = g
        g = self.u0.backward(g)
        g = self.c1.backward(g)
        ds1 = g
        g = self.u1.backward(g)
        g = self.c2.backward(g)
        ds2 = g
        g = self.u2.backward(g)
        g = self.b3.backward(g)
        g = self.d2.backward(g)
        g = self.e2.backward(_add_f(g, ds2))
        g = self.d1.backward(g)
        g = self.e1.backward(_add_f(g, ds1))
        g = self.d0.backward(g)
        g = self.e0b.backward(_add_f(g, ds0))
        self.e0a.backward(g, need_dx=False)

    def step(self, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
        self.step_count += 1
        for _, P in self.named_params():
            _call("afx_nn_adam", _p(P.w), _p(P.g), _p(P.m), _p(P.v), _p(P.wb), P.w.numel(), lr, b1, b2, eps,
                  self.step_count, _stream())

    def loss(self, out, target, mask=None, count=None):
        """Masked MSE over the 6 output channels (mask: uint8 per row, `count` rows set);
        returns (loss tensor [1], grad [V][32])."""
        torch = _torch()
        V = out.shape[0]
        grad = torch.empty_like(out)
        part = torch.empty((V * 32 + 255) // 256, dtype=torch.float32, device=out.device)
        loss = torch.empty(1, dtype=torch.float32, device=out.device)
        cnt = V if mask is None else int(count)
        _call("afx_nn_mse", _p(out), 32, _p(target), _p(mask), V, self.NOUT, 1.0 / max(cnt * self.NOUT, 1), _p(grad),
              _p(part), _p(loss), _stream())
        return loss, grad

    def state(self) -> dict:
        return {name: P.w.detach().cpu().numpy().copy() for name, P in self.named_params()}

    def load_state(self, st: dict):
        for name, P in self.named_params():
            P.w.copy_(_torch().from_numpy(np.asarray(st[name], np.float32)))
            P.refresh()


# ----------------------------------------------------------------------------------------------
# cell grid <-> edges of a structured Kuhn box (mesh.tide_case)

class GridMaps:
    """Device maps between the cell grid (nz, ny, nx) and the edge values of a box case."""

    def __init__(self, case, device="cuda"):
        torch = _torch()
        dev = torch.device(device)
        nline = case.nline
        ncell = case.nx * case.ny * case.nz
        e = np.abs(case.n6line).astype(np.int64) - 1          # (ntet, 6)
        cell = np.repeat(case.cell_of_tet.astype(np.int64)[:, None], 6, axis=1)
        pairs = np.unique(np.stack([cell.ravel(), e.ravel()], 1), axis=0)   # unique (cell, edge)
        # cell -> edges
        order = np.lexsort((pairs[:, 1], pairs[:, 0]))
        pc = pairs[order]
        cptr = np.zeros(ncell + 1, np.int32)
        np.add.at(cptr, pc[:, 0] + 1, 1)
        cptr = np.cumsum(cptr).astype(np.int32)
        # edge -> cells
        order = np.lexsort((pairs[:, 0], pairs[:, 1]))
        pe = pairs[order]
        eptr = np.zeros(nline + 1, np.int32)
        np.add.at(eptr, pe[:, 1] + 1, 1)
        eptr = np.cumsum(eptr).astype(np.int32)
        ln = case.line.astype(np.int64) - 1
        evec = (case.xyz[ln[:, 1]] - case.xyz[ln[:, 0]]).astype(np.float64)
        # edges every incident cell of which is air (group 1): sigma_air leaves their gradient
        # components almost undetermined, so a prediction there only adds error
        cell_air = np.zeros(ncell, bool)
        cell_air[case.cell_of_tet[case.group == 1]] = True
        not_air = np.zeros(nline, np.int64)
        np.add.at(not_air, pe[:, 1], (~cell_air[pe[:, 0]]).astype(np.int64))
        self.air_only = torch.from_numpy((not_air == 0).astype(np.uint8)).to(dev)
        self.ncell, self.nline = ncell, nline
        self.dims = (case.nz, case.ny, case.nx)
        self.cptr = torch.from_numpy(cptr).to(dev)
        self.cedge = torch.from_numpy(pc[:, 1].astype(np.int32)).to(dev)
        self.eptr = torch.from_numpy(eptr).to(dev)
        self.ecell = torch.from_numpy(pe[:, 0].astype(np.int32)).to(dev)
        self.evec = torch.from_numpy(evec.ravel()).to(dev)

    def cells_to_edges(self, F, scale, ldf=32, zero=None, out=None):
        """F: FP32 [ncell][ldf] (channels 0..5) -> complex128 edge values [nline] (zero where the
        uint8 mask `zero` is set)."""
        torch = _torch()
        x = out if out is not None else torch.empty(self.nline, dtype=torch.complex128, device=F.device)
        _call("afx_nn_cells_to_edges", _p(self.eptr), _p(self.ecell), _p(self.evec), _p(F), ldf, float(scale),
              _p(zero), self.nline, _p(x), _stream())
        return x

    def edges_to_cells(self, x, scale, ldf=6):
        """complex128 edge values [nline] -> FP32 [ncell][ldf] least-squares cell fields / scale."""
        torch = _torch()
        F = torch.zeros((self.ncell, ldf), dtype=torch.float32, device=x.device)
        _call("afx_nn_edges_to_cells", _p(self.cptr), _p(self.cedge), _p(self.evec), _p(x), float(scale),
              self.ncell, _p(F), ldf, _stream())
        return F


def cell_features(case, cin=8) -> np.ndarray:
    """Network input per cell (channels-last, (nz*ny*nx, cin) float32):
    0: log10 sigma (air, sea, crust) scaled as (log10 sigma + 2) / 3;
    1..6: Re/Im of the tidal source sigma_sea (v x F) averaged over the cell's nodes (sea cells),
          scaled by its largest magnitude; the rest zero."""
    nz, ny, nx = case.nz, case.ny, case.nx
    ncell = nx * ny * nz
    feat = np.zeros((ncell, cin), np.float32)
    lsig = np.empty(ncell)
    st = case.sigma_tet()
    lsig[case.cell_of_tet] = np.log10(st)
    feat[:, 0] = (lsig + 2.0) / 3.0
    vf = np.cross(case.vxyz, case.fxyz.astype(np.complex128))   # (node, 3)
    n4 = case.n4.astype(np.int64) - 1
    sea = case.group == 2
    J = np.zeros((ncell, 3), np.complex128)
    cnt = np.zeros(ncell)
    for c in range(4):
        np.add.at(J, case.cell_of_tet[sea], vf[n4[sea, c]])
        np.add.at(cnt, case.cell_of_tet[sea], 1.0)
    J[cnt > 0] /= cnt[cnt > 0, None]
    J *= case.sigma_sea
    m = np.abs(J).max()
    if m > 0:
        J /= m
    feat[:, 1:4] = J.real
    feat[:, 4:7] = J.imag
    return feat
