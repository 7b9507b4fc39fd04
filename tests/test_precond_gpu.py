"""The learned preconditioner's kernels and solver on the GPU.

* Convolutions (BF16 operands, FP32 accumulation on the matrix cores) against
  torch's FP32 convolution of the same BF16-rounded operands: forward to 1e-5
  of the output's scale, weight and data gradients the same (the reference
  gradients computed in FP32 on the CPU).
* Pooling and upsampling and their gradients exactly against their
  definitions.
* The whole U-Net's weight gradients against torch autograd of the same
  network in FP32 on the CPU (cosine above 0.99 and norms within 5 % per
  layer: the device rounds activations and gradients to BF16 between layers;
  a wrong offset, transpose or routing gives cosines far from one).
* Flexible CG with the identity preconditioner takes the same steps as the
  reference's plain CG (core/src/clean.rs:263-306) on the exact Gram operator.
* Training on the gridder's residuals lowers the loss, and flexible CG with
  the trained preconditioner converges, judged by the exact Gram residual;
  its Gram applications are reported beside plain CG's.
"""

import math

import numpy as np
import pytest
from conftest import ref_case

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
F = torch.nn.functional


def bf16_bits(t):
    return t.to(torch.bfloat16).view(torch.int16)


def from_bits(b):
    return b.view(torch.bfloat16).to(torch.float32)


def nchw(x_nhwc, B, H, W, C):
    return x_nhwc.reshape(B, H, W, C).permute(0, 3, 1, 2)


def torch_weight(w, ks):
    # [cout][k*k][cin] -> [cout][cin][k][k]
    co, _, ci = w.shape
    return w.reshape(co, ks, ks, ci).permute(0, 3, 1, 2)


@pytest.mark.parametrize("ks,cin,cout", [(3, 32, 48), (3, 16, 16), (1, 32, 16)])
def test_conv_forward_and_gradients(ks, cin, cout):
    from rang import precond as P

    lib = P._lib()
    st = P._stream()
    g = torch.Generator().manual_seed(ks * 100 + cin + cout)
    B, H, W = 2, 12, 20
    ldx, xoff = cin + 16, 16  # a strided slice, as the concatenations use
    x = torch.randn(B * H * W, ldx, generator=g)
    w = torch.randn(cout, ks * ks, cin, generator=g) / math.sqrt(ks * ks * cin)
    xb, wb = bf16_bits(x).cuda(), bf16_bits(w).cuda()
    y = torch.zeros(B * H * W, cout + 8, device="cuda")
    a = torch.zeros(B * H * W, cout + 8, dtype=torch.int16, device="cuda")
    P._ok(lib.rang_amd_nn_conv_fwd(P._p(xb), ldx, xoff, P._p(wb), P._p(y), P._p(a), cout + 8, 8, B, H, W, cin, cout,
                                   ks, 1, st))
    xr = nchw(from_bits(xb.cpu())[:, xoff:xoff + cin].contiguous(), B, H, W, cin)
    wr = torch_weight(from_bits(wb.cpu()), ks)
    ref = F.conv2d(xr, wr, padding=ks // 2)  # FP32 on the CPU
    got = nchw(y.cpu()[:, 8:8 + cout].contiguous(), B, H, W, cout)
    scale = float(ref.abs().max())
    assert float((got - ref).abs().max()) <= 1e-5 * scale
    act = nchw(from_bits(a.cpu())[:, 8:8 + cout].contiguous(), B, H, W, cout)
    assert float((act - ref.clamp(min=0)).abs().max()) <= 1e-2 * scale  # BF16 rounding of relu(y)
    # gradients for an output gradient dy
    dy = torch.randn(B * H * W, cout, generator=g)
    dyb = bf16_bits(dy).cuda()
    dw = torch.zeros(cout, ks * ks, cin, device="cuda")
    P._ok(lib.rang_amd_nn_conv_wgrad(P._p(xb), ldx, xoff, P._p(dyb), cout, 0, P._p(dw), B, H, W, cin, cout, ks, st))
    dyr = nchw(from_bits(dyb.cpu()), B, H, W, cout)
    ref_dw = torch.nn.grad.conv2d_weight(xr, wr.shape, dyr, padding=ks // 2)
    got_dw = torch_weight(dw.cpu(), ks)
    assert float((got_dw - ref_dw).abs().max()) <= 1e-5 * float(ref_dw.abs().max())
    wf = w.cuda().contiguous()
    wt = torch.zeros(cin, ks * ks, cout, dtype=torch.int16, device="cuda")
    P._ok(lib.rang_amd_nn_flip(P._p(wf), P._p(wt), cin, cout, ks, st))
    dx = torch.zeros(B * H * W, cin, device="cuda")
    P._ok(lib.rang_amd_nn_conv_fwd(P._p(dyb), cout, 0, P._p(wt), P._p(dx), None, cin, 0, B, H, W, cout, cin, ks, 0,
                                   st))
    ref_dx = torch.nn.grad.conv2d_input(xr.shape, torch_weight(from_bits(bf16_bits(w)), ks), dyr, padding=ks // 2)
    got_dx = nchw(dx.cpu(), B, H, W, cin)
    assert float((got_dx - ref_dx).abs().max()) <= 1e-5 * float(ref_dx.abs().max())


def test_pool_and_upsample_with_gradients():
    from rang import precond as P

    lib, st = P._lib(), P._stream()
    g = torch.Generator().manual_seed(3)
    B, H, W, C = 2, 8, 6, 16
    x = torch.randn(B * H * W, C, generator=g)
    xb = bf16_bits(x).cuda()
    y = torch.zeros(B * H * W // 4, C, dtype=torch.int16, device="cuda")
    P._ok(lib.rang_amd_nn_pool(P._p(xb), C, 0, P._p(y), C, 0, B, H, W, C, st))
    xr = nchw(from_bits(xb.cpu()), B, H, W, C)
    ref = F.avg_pool2d(xr, 2)
    assert float((nchw(from_bits(y.cpu()), B, H // 2, W // 2, C) - ref).abs().max()) <= 1e-2
    up = torch.zeros(B * H * W, C, dtype=torch.int16, device="cuda")
    P._ok(lib.rang_amd_nn_up(P._p(y), C, 0, P._p(up), C, 0, B, H, W, C, st))
    assert torch.equal(nchw(up.cpu(), B, H, W, C),
                       F.interpolate(nchw(y.cpu(), B, H // 2, W // 2, C).float(), scale_factor=2).to(torch.int16))
    gy = torch.randn(B * H * W // 4, C, generator=g).cuda()
    gx = torch.ones(B * H * W, C, device="cuda")
    P._ok(lib.rang_amd_nn_pool_bwd(P._p(gy), C, 0, P._p(gx), C, 0, B, H, W, C, st))
    ref_gx = 1.0 + F.interpolate(nchw(gy.cpu(), B, H // 2, W // 2, C), scale_factor=2) / 4.0
    assert float((nchw(gx.cpu(), B, H, W, C) - ref_gx).abs().max()) <= 1e-6
    gf = torch.randn(B * H * W, C, generator=g).cuda()
    gc = torch.zeros(B * H * W // 4, C, device="cuda")
    P._ok(lib.rang_amd_nn_up_bwd(P._p(gf), C, 0, P._p(gc), C, 0, B, H, W, C, st))
    ref_gc = 4.0 * F.avg_pool2d(nchw(gf.cpu(), B, H, W, C), 2)
    assert float((nchw(gc.cpu(), B, H // 2, W // 2, C) - ref_gc).abs().max()) <= 1e-5


def torch_unet(net, x_in):
    """The same network in FP32 on the CPU with autograd (BF16-rounded weights)."""
    from rang import precond as P

    ws = []

    def conv(k, x, relu):
        w = torch_weight(from_bits(bf16_bits(k.w.detach().cpu())), k.ks).clone().requires_grad_(True)
        ws.append((k, w))
        y = F.conv2d(x, w, padding=k.ks // 2)
        return F.relu(y) if relu else y

    L = net.levels
    skips = []
    h = x_in
    for l in range(L):
        ka, kb = net.enc[l]
        h = conv(kb, conv(ka, h, True), True)
        skips.append(h)
        if l < L - 1:
            h = F.avg_pool2d(h, 2)
    d = skips[L - 1]
    for l in range(L - 2, -1, -1):
        ka, kb = net.dec[l]
        cat = torch.cat([F.interpolate(d, scale_factor=2), skips[l]], dim=1)
        d = conv(kb, conv(ka, cat, True), True)
    out = conv(net.out, d, False)
    assert out.shape[1] == P.COUT
    return out, ws


def test_unet_gradients_match_autograd():
    from rang import precond as P

    torch.manual_seed(0)
    B, H, W = 2, 16, 24
    net = P.UNet(H, W, batch=B, levels=3, base=16, seed=4)
    g = torch.Generator().manual_seed(5)
    x = torch.randn(B * H * W, P.CIN, generator=g)
    x[:, 2:] = 0.0
    net.buf[0]["in"].copy_(bf16_bits(x).cuda())
    y = net.forward(B)
    t = torch.randn(B, H * W, generator=g)
    inv = torch.tensor([1.0 / B, 0.5 / B], dtype=torch.float64)
    loss = torch.zeros(1, dtype=torch.float64, device="cuda")
    tt, invd = t.cuda().contiguous(), inv.cuda()
    P._ok(P._lib().rang_amd_nn_loss(P._p(y), P.COUT, 0, P._p(tt), P._p(invd), P._p(net.g_y), P.COUT, B, H * W,
                                     P._p(loss), P._stream()))
    net.zero_grad()
    net.backward(B)
    xr = nchw(from_bits(bf16_bits(x)), B, H, W, P.CIN)
    out, ws = torch_unet(net, xr)
    ref_loss = sum(float(inv[n]) * 0.5 * ((out[n, 0].reshape(-1) - t[n]) ** 2).sum() for n in range(B))
    assert abs(float(loss[0]) - float(ref_loss.detach())) <= 2e-2 * float(ref_loss.detach())
    ref_loss.backward()
    for k, w in ws:
        got = torch_weight(k.g.cpu(), k.ks).reshape(-1).double()
        ref = w.grad.reshape(-1).double()
        cos = float(torch.dot(got, ref) / (got.norm() * ref.norm()))
        ratio = float(got.norm() / ref.norm())
        print(f"[precond] layer {k.cin}->{k.cout} k{k.ks}: cosine {cos:.5f}, norm ratio {ratio:.4f}")
        assert cos > 0.99, (k.cin, k.cout, k.ks, cos)
        assert abs(ratio - 1.0) < 5e-2


@pytest.fixture(scope="module")
def small_problem():
    # the array and sky recipe of the benchmark workloads, small: 2016
    # baselines x 16 times x 4 channels, 64^2 at 4 arcsec
    from rang import synth
    from rang.measurement import WgridMeasurement

    c = synth.imaging_case_device(16, 4, 1.30e9, 1.45e9, 64, 4.0, torch.device("cuda:0"))
    m = WgridMeasurement(c["uvw"], c["freq"], c["vis"], c["wgt"].double(), 64, c["pixsize"], 1e-7)
    rhs, _ = m.residual(torch.zeros((64, 64), dtype=torch.float64, device="cuda"))
    return m, rhs


def test_flexible_cg_with_the_identity_is_cg(small_problem):
    from rang import precond as P
    from rang.measurement import conjugate_gradients

    m, rhs = small_problem
    x_cg, used_cg, rel_cg = conjugate_gradients(m, rhs, 40, 1e-6)
    for trunc in (1, 4):
        x_f, used_f, true_rel = P.preconditioned_cg(m, lambda r: r.clone(), rhs, 40, 1e-6, truncation=trunc)
        assert abs(used_f - used_cg) <= 1
        assert float((x_f - x_cg).abs().max()) <= 1e-4 * float(x_cg.abs().max())
        # the exact Gram residual of the result
        assert true_rel <= max(rel_cg, 1e-6) * 1.5


def test_training_and_the_learned_solve(small_problem):
    # Training on the gridder's own residuals lowers the preconditioner's
    # residual loss, and flexible CG with the trained network converges,
    # judged by the exact Gram residual. The Gram applications of both
    # solves are reported; the learned one is not (yet) the fewer.
    from rang import precond as P
    from rang.measurement import conjugate_gradients

    m, rhs = small_problem
    net = P.UNet(64, 64, batch=8, levels=3, base=16, seed=2)
    tr = P.Trainer(m, net, seed=11)
    n = tr.collect(rhs_count=8, iterations=60, keep=(0, 5, 10, 20, 40, 59))
    assert n > 0
    losses = tr.train(120, lr=2e-3, lr_end=2e-4)
    first, last = np.mean(losses[:10]), np.mean(losses[-10:])
    print(f"[precond] loss {first:.4f} -> {last:.4f} over {len(losses)} steps, {tr.gram_applies} Gram applies")
    assert last < 0.7 * first
    M = P.Preconditioner(m, net)
    tol = 1e-2
    _, used_cg, rel_cg = conjugate_gradients(m, rhs, 2000, tol)
    stats = {}
    _, used_p, true_rel = P.preconditioned_cg(m, M, rhs, 2000, tol, stats=stats)
    print(f"[precond] Gram applies to {tol:g}: CG {used_cg} (rel {rel_cg:.2e}), learned {used_p} "
          f"(exact residual {true_rel:.2e}, fallbacks {stats['fallbacks']})")
    assert true_rel <= 1.5 * tol
    assert M.applies == used_p
