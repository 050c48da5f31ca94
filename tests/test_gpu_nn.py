"""Surrogate kernels (include/afx_nn.h) against PyTorch-CPU float32 on the same bf16 values.

Each GEMM-lowered layer is compared with torch's own conv3d / conv_transpose3d / linear and
their autograd gradients, computed on the CPU in float32 from exactly the bf16 inputs and weights
the GPU path uses (so the only difference is the FP32 summation order). The full network is then
checked end to end against a torch-CPU model of the same architecture whose activations are
rounded to bf16 where the GPU path rounds them."""
import math

import numpy as np
import pytest

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu


def _bf(x):
    """float32 tensor rounded to bf16 values (kept as float32)."""
    return x.to(torch.bfloat16).to(torch.float32)


def _rel(a, b):
    a = a.double()
    b = b.double()
    return float((a - b).norm() / max(float(b.norm()), 1e-30))


def _dev():
    return torch.device("cuda:0")


def test_gemm_nt_matches_cpu():
    from actfemtide_amd import nn
    g = torch.Generator().manual_seed(0)
    for (M, N, K, splits, relu) in [(300, 70, 96, 1, False), (129, 32, 48, 1, True), (64, 200, 1024, 5, False),
                                    (1000, 6, 64, 1, False)]:
        A = _bf(torch.randn(M, K, generator=g))
        B = _bf(torch.randn(N, K, generator=g))
        bias = torch.randn(N, generator=g)
        ref = A @ B.T
        dA = A.to(torch.bfloat16).to(_dev())
        dB = B.to(torch.bfloat16).to(_dev())
        C = torch.zeros(M, N, dtype=torch.float32, device=_dev())
        if splits > 1:
            nn.gemm_nt(dA, dB, M, N, K, C_out=C, splits=splits)
        else:
            ref = ref + bias
            if relu:
                ref = torch.relu(ref)
            nn.gemm_nt(dA, dB, M, N, K, C_out=C, bias=bias.to(_dev()), relu=relu)
        torch.cuda.synchronize()
        err = _rel(C.cpu(), ref)
        print(f"gemm M={M} N={N} K={K} splits={splits}: rel err {err:.2e}")
        assert err < 1e-5


def test_gemm_nt_automatic_split_k_with_epilogue():
    """A deep-level shape (few output tiles, long K): the wrapper splits K by itself, and the
    split-order sum then applies the bias, the ReLU and both outputs (FP32 and bf16)."""
    from actfemtide_amd import nn
    g = torch.Generator().manual_seed(1)
    M, N, K = 192, 256, 6912
    A = _bf(torch.randn(M, K, generator=g))
    B = _bf(torch.randn(N, K, generator=g))
    bias = torch.randn(N, generator=g)
    ref = torch.relu(A @ B.T + bias)
    C = torch.zeros(M, N, dtype=torch.float32, device=_dev())
    Cb = torch.zeros(M, N, dtype=torch.bfloat16, device=_dev())
    nn.gemm_nt(A.to(torch.bfloat16).to(_dev()), B.to(torch.bfloat16).to(_dev()), M, N, K, C_out=C,
               Cb=Cb.view(torch.int16), bias=bias.to(_dev()), relu=True)
    torch.cuda.synchronize()
    err, errb = _rel(C.cpu(), ref), _rel(Cb.float().cpu(), ref)
    print(f"gemm M={M} N={N} K={K} automatic split: rel err {err:.2e}, bf16 output {errb:.2e}")
    assert err < 1e-5 and errb < 1e-2


def _conv_w(layer):
    """torch conv3d weight [co][ci][3][3][3] from Conv3's W [co][27*ci padded] (tap = (dz*3+dy)*3+dx)."""
    co, ci = layer.co, layer.ci
    w = layer.W.wb.float().cpu()[:, :27 * ci].reshape(co, 27, ci)
    return w.permute(0, 2, 1).reshape(co, ci, 3, 3, 3).contiguous()


def _ncdhw(x, dims):
    nb, D, H, W = dims
    return x.reshape(nb, D, H, W, -1).permute(0, 4, 1, 2, 3).contiguous()


def _ndhwc(x):
    return x.permute(0, 2, 3, 4, 1).reshape(-1, x.shape[1]).contiguous()


def test_conv3_forward_backward():
    from actfemtide_amd import nn
    rng = np.random.default_rng(3)
    dims = (2, 4, 6, 8)
    V = int(np.prod(dims))
    L = nn.Conv3(8, 16, rng, _dev())
    L.b.w.copy_(torch.randn(16) * 0.1)
    x = _bf(torch.randn(V, 8))
    y = L.forward(x.to(torch.bfloat16).to(_dev()), dims)
    dy = _bf(torch.randn(V, 16))
    dx = L.backward(dy.to(_dev()))
    torch.cuda.synchronize()
    # torch reference from the same bf16 weights / inputs
    xr = _ncdhw(x, dims).requires_grad_(True)
    w = _conv_w(L).requires_grad_(True)
    b = L.b.w.cpu().clone().requires_grad_(True)
    pre = torch.nn.functional.conv3d(xr, w, b, padding=1)
    out = torch.relu(pre)
    out.backward(_ncdhw(dy, dims))
    err_y = _rel(y.float().cpu(), _bf(_ndhwc(out.detach())))
    err_dx = _rel(dx.cpu(), _ndhwc(xr.grad))
    gw = L.W.g.cpu()[:, :27 * 8].reshape(16, 27, 8).permute(0, 2, 1).reshape(16, 8, 3, 3, 3)
    err_dw = _rel(gw, w.grad)
    err_db = _rel(L.b.g.cpu(), b.grad)
    print(f"conv3: y {err_y:.2e} dx {err_dx:.2e} dW {err_dw:.2e} db {err_db:.2e}")
    assert err_y < 5e-3 and err_dx < 1e-4 and err_dw < 1e-4 and err_db < 1e-5
    # padded weight columns get no gradient
    assert float(L.W.g[:, 27 * 8:].abs().max()) == 0.0


def test_down_up_head_layers():
    from actfemtide_amd import nn
    rng = np.random.default_rng(4)
    dims = (1, 4, 4, 8)
    V = int(np.prod(dims))
    # stride-2 down: W [co][8*ci], k = s*ci + c, s = (dz*2 + dy)*2 + dx
    Ld = nn.Down(16, 32, rng, _dev())
    x = _bf(torch.randn(V, 16))
    y, g2 = Ld.forward(x.to(torch.bfloat16).to(_dev()), dims)
    Vo = V // 8
    dy = _bf(torch.randn(Vo, 32))
    dx = Ld.backward(dy.to(_dev()))
    xr = _ncdhw(x, dims).requires_grad_(True)
    w = Ld.W.wb.float().cpu().reshape(32, 2, 2, 2, 16).permute(0, 4, 1, 2, 3).contiguous().requires_grad_(True)
    b = Ld.b.w.cpu().clone().requires_grad_(True)
    out = torch.relu(torch.nn.functional.conv3d(xr, w, b, stride=2))
    out.backward(_ncdhw(dy, g2))
    gw = Ld.W.g.cpu().reshape(32, 2, 2, 2, 16).permute(0, 4, 1, 2, 3)
    errs = (_rel(y.float().cpu(), _bf(_ndhwc(out.detach()))), _rel(dx.cpu(), _ndhwc(xr.grad)),
            _rel(gw, w.grad), _rel(Ld.b.g.cpu(), b.grad))
    print("down: y %.2e dx %.2e dW %.2e db %.2e" % errs)
    assert errs[0] < 5e-3 and max(errs[1:]) < 1e-4
    # transposed stride-2 up: W [8*co][ci], row n = s*co + c
    Lu = nn.Up(32, 16, rng, _dev())
    Lu.b.w.copy_(torch.randn(16) * 0.1)
    Lu.b.refresh()
    xc = _bf(torch.randn(Vo, 32))
    yu, gf = Lu.forward(xc.to(torch.bfloat16).to(_dev()), g2)
    dyu = _bf(torch.randn(V, 16))
    dxu = Lu.backward(dyu.to(_dev()))
    xr = _ncdhw(xc, g2).requires_grad_(True)
    w = Lu.W.wb.float().cpu().reshape(2, 2, 2, 16, 32).permute(4, 3, 0, 1, 2).contiguous().requires_grad_(True)
    b = Lu.b.w.cpu().clone().requires_grad_(True)
    out = torch.relu(torch.nn.functional.conv_transpose3d(xr, w, b, stride=2))
    out.backward(_ncdhw(dyu, dims))
    gw = Lu.W.g.cpu().reshape(2, 2, 2, 16, 32).permute(4, 3, 0, 1, 2)
    errs = (_rel(yu.float().cpu(), _bf(_ndhwc(out.detach()))), _rel(dxu.cpu(), _ndhwc(xr.grad)),
            _rel(gw, w.grad), _rel(Lu.b.g.cpu(), b.grad))
    print("up: y %.2e dx %.2e dW %.2e db %.2e" % errs)
    assert errs[0] < 5e-3 and max(errs[1:]) < 1e-4
    # 1x1 head, linear, FP32 out
    Lh = nn.Head(16, rng, _dev())
    xh = _bf(torch.randn(V, 16))
    o = Lh.forward(xh.to(torch.bfloat16).to(_dev()), dims)
    do = torch.zeros(V, 32)
    do[:, :6] = _bf(torch.randn(V, 6))
    dxh = Lh.backward(do.to(_dev()))
    xr = xh.clone().requires_grad_(True)
    w = Lh.W.wb.float().cpu()[:6].clone().requires_grad_(True)
    b = Lh.b.w.cpu()[:6].clone().requires_grad_(True)
    out = xr @ w.T + b
    out.backward(do[:, :6])
    errs = (_rel(o.cpu()[:, :6], out.detach()), _rel(dxh.cpu(), xr.grad), _rel(Lh.W.g.cpu()[:6], w.grad),
            _rel(Lh.b.g.cpu()[:6], b.grad))
    print("head: y %.2e dx %.2e dW %.2e db %.2e" % errs)
    assert max(errs) < 1e-4
    assert float(Lh.W.g[6:].abs().max()) == 0.0


class _RefNet(torch.nn.Module):
    """torch-CPU model of nn.Surrogate from its bf16 weights; activations rounded to bf16
    (straight-through) where the GPU path stores them in bf16."""

    def __init__(self, s):
        super().__init__()
        self.s = s
        P = {}
        for L in ["e0a", "e0b", "e1", "e2", "b3", "c2", "c1", "c0"]:
            P[L + ".W"] = _conv_w(getattr(s, L))
        for L in ["d0", "d1", "d2"]:
            m = getattr(s, L)
            P[L + ".W"] = m.W.wb.float().cpu().reshape(m.co, 2, 2, 2, m.ci).permute(0, 4, 1, 2, 3).contiguous()
        for L in ["u2", "u1", "u0"]:
            m = getattr(s, L)
            P[L + ".W"] = m.W.wb.float().cpu().reshape(2, 2, 2, m.co, m.ci).permute(4, 3, 0, 1, 2).contiguous()
        P["head.W"] = s.head.W.wb.float().cpu()[:6].clone()
        for L in s.layers:
            b = getattr(s, L).b.w.cpu().clone()
            P[L + ".b"] = b[:6] if L == "head" else b
        self.P = {k: torch.nn.Parameter(v) for k, v in P.items()}

    def forward(self, x):
        P = self.P
        F = torch.nn.functional

        def r(t):
            return t + (_bf(t) - t).detach()

        def c3(t, n):
            return r(torch.relu(F.conv3d(t, P[n + ".W"], P[n + ".b"], padding=1)))

        def dn(t, n):
            return r(torch.relu(F.conv3d(t, P[n + ".W"], P[n + ".b"], stride=2)))

        def up(t, n):
            return r(torch.relu(F.conv_transpose3d(t, P[n + ".W"], P[n + ".b"], stride=2)))

        s0 = c3(c3(x, "e0a"), "e0b")
        s1 = c3(dn(s0, "d0"), "e1")
        s2 = c3(dn(s1, "d1"), "e2")
        a = c3(dn(s2, "d2"), "b3")
        a = c3(r(up(a, "u2") + s2), "c2")
        a = c3(r(up(a, "u1") + s1), "c1")
        a = c3(r(up(a, "u0") + s0), "c0")
        y = a.permute(0, 2, 3, 4, 1).reshape(-1, a.shape[1])
        return y @ P["head.W"].T + P["head.b"]


def test_surrogate_forward_backward_matches_torch():
    from actfemtide_amd import nn
    dims = (2, 8, 8, 16)
    V = int(np.prod(dims))
    s = nn.Surrogate(base=16, cin=8, seed=5, device="cuda")
    g = torch.Generator().manual_seed(1)
    for _, P in s.named_params():   # non-zero biases exercise every bias path
        if P.w.dim() == 1:
            P.w.copy_(torch.randn(P.w.shape, generator=g) * 0.05)
            P.refresh()
    s.head.b.w[6:] = 0
    s.head.b.refresh()
    x = _bf(torch.randn(V, 8, generator=g))
    out = s.forward(x.to(torch.bfloat16).to(_dev()), dims)
    tgt = torch.randn(V, 6, generator=g)
    loss, grad = s.loss(out, tgt.to(_dev()))
    s.backward(grad)
    torch.cuda.synchronize()
    ref = _RefNet(s)
    ro = ref(_ncdhw(x, dims))
    rl = ((ro - tgt) ** 2).mean()
    rl.backward()
    err_o = _rel(out.cpu()[:, :6], ro.detach())
    print(f"surrogate: out rel err {err_o:.2e}, loss {float(loss.cpu()[0]):.6f} vs {float(rl):.6f}")
    assert err_o < 3e-2
    assert abs(float(loss.cpu()[0]) - float(rl)) < 3e-2 * float(rl)
    worst = 0.0
    for name, P in s.named_params():
        gr = ref.P[name].grad
        gg = P.g.cpu()
        L = name.split(".")[0]
        if name.endswith(".W"):
            m = getattr(s, L)
            if L in ("e0a", "e0b", "e1", "e2", "b3", "c2", "c1", "c0"):
                gg = gg[:, :27 * m.ci].reshape(m.co, 27, m.ci).permute(0, 2, 1).reshape(m.co, m.ci, 3, 3, 3)
            elif L.startswith("d"):
                gg = gg.reshape(m.co, 2, 2, 2, m.ci).permute(0, 4, 1, 2, 3)
            elif L.startswith("u"):
                gg = gg.reshape(2, 2, 2, m.co, m.ci).permute(4, 3, 0, 1, 2)
            else:
                gg = gg[:6]
        elif L == "head":
            gg = gg[:6]
        e = _rel(gg, gr)
        cos = float((gg.double() * gr.double()).sum() / (gg.double().norm() * gr.double().norm() + 1e-300))
        worst = max(worst, e)
        print(f"  {name:8s} rel {e:.2e} cos {cos:.6f}")
        assert e < 8e-2 and cos > 0.996, name
    print(f"worst gradient rel err {worst:.2e}")


def test_adam_step_decreases_loss():
    from actfemtide_amd import nn
    dims = (1, 8, 8, 8)
    V = int(np.prod(dims))
    s = nn.Surrogate(base=16, cin=8, seed=7, device="cuda")
    g = torch.Generator().manual_seed(2)
    x = _bf(torch.randn(V, 8, generator=g)).to(torch.bfloat16).to(_dev())
    tgt = (torch.randn(V, 6, generator=g) * 0.5).to(_dev())
    losses = []
    for it in range(60):
        out = s.forward(x, dims)
        loss, grad = s.loss(out, tgt)
        s.backward(grad)
        s.step(lr=3e-3)
        losses.append(float(loss.cpu()[0]))
    print("losses", losses[0], losses[-1])
    assert losses[-1] < 0.5 * losses[0]


def test_cell_edge_maps_reproduce_constant_fields():
    """Whitney-form exactness on the cell grid: a constant complex field F gives edge values
    F . t_e, which every cell's least-squares fit returns exactly, and averaging any cell's
    constant field back onto its edges reproduces the edge values."""
    from actfemtide_amd import mesh, nn
    case = mesh.tide_case(8, 8, 8, seed=3)
    gm = nn.GridMaps(case, device="cuda")
    Fc = np.array([1.5 - 0.5j, -2.0 + 0.25j, 0.75 + 1.0j])
    ln = case.line.astype(np.int64) - 1
    t = case.xyz[ln[:, 1]] - case.xyz[ln[:, 0]]
    xe = t @ Fc
    F = gm.edges_to_cells(torch.from_numpy(xe).to(_dev()), scale=2.0)
    want = np.concatenate([Fc.real, Fc.imag]) / 2.0
    got = F.cpu().numpy()
    assert np.abs(got - want[None, :]).max() < 1e-6 * np.abs(want).max()
    F32 = torch.zeros((gm.ncell, 32), dtype=torch.float32)
    F32[:, :6] = torch.from_numpy(np.tile(want.astype(np.float32), (gm.ncell, 1)))
    xb = gm.cells_to_edges(F32.to(_dev()), scale=2.0).cpu().numpy()
    assert np.abs(xb - xe).max() < 1e-6 * np.abs(xe).max()
