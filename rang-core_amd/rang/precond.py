"""A learned preconditioner for the normal equations A^H W A x = A^H W v of
the exact gridder: a small U-Net whose convolutions run in BF16 with FP32
accumulation on the matrix cores (hand-written kernels behind
include/rang_core_precond.h), trained on images the gridder itself produces,
and used inside flexible conjugate gradients. Plain CG
(`rang.measurement.conjugate_gradients`, core/src/clean.rs:263-306) stays the
fallback, and the exact Gram operator (core/src/clean_certified.rs:233-236)
judges every residual. The reference has no learned preconditioner.

    net = UNet(side, side, batch=8, levels=4, base=32)
    trainer = Trainer(measurement, net)
    trainer.train(steps=300)
    x, applies, rel = preconditioned_cg(measurement, net, rhs, 200, 1e-4)

The network maps a residual r (with the plan's PSF as a second input
channel) to an estimate of G^{-1} r. It has no biases, so it is positively
homogeneous; the preconditioner applies it to r / |r| and -r / |r| and takes
half the difference, times |r|, which makes it odd and homogeneous of degree
one, as a preconditioner of a linear system should be. Torch only holds
device memory and runs the calls on its current stream.
"""

import ctypes
import math

import numpy as np
import torch

from . import _core
from .device import _stream

CIN = 16  # the padded input channels: residual, PSF, zeros
COUT = 16  # the padded output channels: channel 0 is the estimate


def _lib():
    return _core.library()


def _p(t):
    return None if t is None else ctypes.c_void_p(t.data_ptr())


def _ok(status):
    if status == 0:
        return
    buf = ctypes.create_string_buffer(512)
    n = _lib().rang_amd_nn_last_error(buf, 511)
    raise RuntimeError(f"preconditioner kernel failed ({status}): {buf.raw[:n].decode(errors='replace')}")


class _Conv:
    """One convolution: FP32 master weights [cout][k*k][cin], their BF16 copy,
    the flipped BF16 weights of the data gradient, gradient and Adam moments."""

    def __init__(self, cin, cout, ks, relu, rng, dev, gain=1.0):
        self.cin, self.cout, self.ks, self.relu = cin, cout, ks, relu
        fan = ks * ks * cin
        w = rng.standard_normal((cout, ks * ks, cin)) * (gain * math.sqrt(2.0 / fan))
        self.w = torch.from_numpy(w.astype(np.float32)).to(dev).contiguous()
        self.wb = torch.empty(self.w.shape, dtype=torch.int16, device=dev)
        self.wt = torch.empty((cin, ks * ks, cout), dtype=torch.int16, device=dev)
        self.g = torch.zeros_like(self.w)
        self.m = torch.zeros_like(self.w)
        self.v = torch.zeros_like(self.w)
        self.refresh()

    def refresh(self):
        st = _stream()
        _ok(_lib().rang_amd_nn_to_bf16(_p(self.w), _p(self.wb), self.w.numel(), st))
        _ok(_lib().rang_amd_nn_flip(_p(self.w), _p(self.wt), self.cin, self.cout, self.ks, st))

    def forward(self, x, ldx, xoff, y, a, ldy, yoff, B, H, W):
        _ok(_lib().rang_amd_nn_conv_fwd(_p(x), ldx, xoff, _p(self.wb), _p(y), _p(a), ldy, yoff, B, H, W, self.cin,
                                        self.cout, self.ks, int(self.relu), _stream()))

    def wgrad(self, x, ldx, xoff, dy, lddy, dyoff, B, H, W):
        _ok(_lib().rang_amd_nn_conv_wgrad(_p(x), ldx, xoff, _p(dy), lddy, dyoff, _p(self.g), B, H, W, self.cin,
                                          self.cout, self.ks, _stream()))

    def dgrad(self, dy, lddy, dyoff, gx, ldgx, gxoff, B, H, W):
        # the input gradient: conv with the flipped weights, channels swapped
        _ok(_lib().rang_amd_nn_conv_fwd(_p(dy), lddy, dyoff, _p(self.wt), _p(gx), None, ldgx, gxoff, B, H, W,
                                        self.cout, self.cin, self.ks, 0, _stream()))

    def adam(self, step, lr, b1=0.9, b2=0.999, eps=1e-8):
        _ok(_lib().rang_amd_nn_adam(_p(self.w), _p(self.g), _p(self.m), _p(self.v), _p(self.wb), self.w.numel(),
                                    lr, b1, b2, eps, step, _stream()))
        _ok(_lib().rang_amd_nn_flip(_p(self.w), _p(self.wt), self.cin, self.cout, self.ks, _stream()))


class UNet:
    """levels x (two 3x3 convolutions with ReLU), 2x2 average pooling down,
    nearest upsampling up with the skip concatenated, a 1x1 output
    convolution; base channels doubling per level; no biases. Buffers for
    `batch` images of H x W (H and W divisible by 2^(levels - 1))."""

    def __init__(self, H, W, *, batch=8, levels=4, base=32, seed=1, device="cuda:0", linear=False):
        if H % (1 << (levels - 1)) or W % (1 << (levels - 1)):
            raise ValueError("image sides must be divisible by 2^(levels - 1)")
        if base % 16:
            raise ValueError("base channels must be a multiple of 16")
        self.H, self.W, self.batch, self.levels, self.base = H, W, batch, levels, base
        self.linear = bool(linear)
        act = not self.linear
        dev = torch.device(device)
        rng = np.random.default_rng(seed)
        L = levels
        c = [base << l for l in range(L)]
        self.c = c
        self.enc = []
        for l in range(L):
            cin = CIN if l == 0 else c[l - 1]
            self.enc.append((_Conv(cin, c[l], 3, act, rng, dev), _Conv(c[l], c[l], 3, act, rng, dev)))
        self.dec = [None] * L
        for l in range(L - 2, -1, -1):
            self.dec[l] = (_Conv(c[l + 1] + c[l], c[l], 3, act, rng, dev), _Conv(c[l], c[l], 3, act, rng, dev))
        self.out = _Conv(c[0], COUT, 1, False, rng, dev, gain=0.1)
        self.convs = [k for pair in self.enc for k in pair] + [k for pair in self.dec if pair for k in pair] + [self.out]
        # activation and gradient buffers per level, for `batch` images
        f32, b16 = torch.float32, torch.int16
        self.buf = []
        for l in range(L):
            P = batch * (H >> l) * (W >> l)
            cl = c[l]
            ldcat = c[l + 1] + cl if l < L - 1 else cl
            b = dict(P=P, ld=ldcat)
            b["in"] = torch.zeros((P, CIN if l == 0 else c[l - 1]), dtype=b16, device=dev)
            b["a_pre"] = torch.empty((P, cl), dtype=f32, device=dev)
            b["a_act"] = torch.empty((P, cl), dtype=b16, device=dev)
            b["b_pre"] = torch.empty((P, ldcat), dtype=f32, device=dev)
            b["cat"] = torch.zeros((P, ldcat), dtype=b16, device=dev)  # [up | skip]; the bottom: its output
            b["g_cat"] = torch.zeros((P, ldcat), dtype=f32, device=dev)
            b["g_a"] = torch.empty((P, cl), dtype=f32, device=dev)
            b["gb16"] = torch.empty((P, max(ldcat, COUT)), dtype=b16, device=dev)
            if l < L - 1:
                b["da_pre"] = torch.empty((P, cl), dtype=f32, device=dev)
                b["da_act"] = torch.empty((P, cl), dtype=b16, device=dev)
                b["db_pre"] = torch.empty((P, cl), dtype=f32, device=dev)
                b["db_act"] = torch.empty((P, cl), dtype=b16, device=dev)
                b["g_d"] = torch.empty((P, cl), dtype=f32, device=dev)
            if l > 0:
                b["g_in"] = torch.empty((P, c[l - 1]), dtype=f32, device=dev)
            self.buf.append(b)
        self.y = torch.empty((batch * H * W, COUT), dtype=f32, device=dev)
        self.g_y = torch.empty((batch * H * W, COUT), dtype=f32, device=dev)
        self.g_bottom = torch.empty((self.buf[L - 1]["P"], c[L - 1]), dtype=f32, device=dev)
        self.step = 0

    def parameters(self):
        return sum(k.w.numel() for k in self.convs)

    def forward_flops(self, B):
        """Multiply-add flops (2 per product) of the convolutions of one
        forward pass over B images, as the matrix cores compute them (the
        padded input and output channels included)."""
        L, c = self.levels, self.c
        total = 0
        for l in range(L):
            P = B * (self.H >> l) * (self.W >> l)
            ka, kb = self.enc[l]
            total += 2 * P * 9 * (ka.cin * ka.cout + kb.cin * kb.cout)
            if self.dec[l]:
                da, db = self.dec[l]
                total += 2 * P * 9 * (da.cin * da.cout + db.cin * db.cout)
        total += 2 * B * self.H * self.W * self.out.cin * self.out.cout
        return total

    # ------------------------------------------------------------- forward
    def forward(self, B):
        """The output (channel 0 of self.y) of the inputs packed in buf[0]['in']."""
        L, c, lib, st = self.levels, self.c, _lib(), _stream()
        for l in range(L):
            b = self.buf[l]
            H, W = self.H >> l, self.W >> l
            ka, kb = self.enc[l]
            ka.forward(b["in"], b["in"].shape[1], 0, b["a_pre"], b["a_act"], c[l], 0, B, H, W)
            skip_off = c[l + 1] if l < L - 1 else 0
            kb.forward(b["a_act"], c[l], 0, b["b_pre"], b["cat"], b["ld"], skip_off, B, H, W)
            if l < L - 1:
                nb = self.buf[l + 1]
                _ok(lib.rang_amd_nn_pool(_p(b["cat"]), b["ld"], skip_off, _p(nb["in"]), c[l], 0, B, H, W, c[l], st))
        for l in range(L - 2, -1, -1):
            b, nb = self.buf[l], self.buf[l + 1]
            H, W = self.H >> l, self.W >> l
            if l + 1 == L - 1:
                src, ldsrc = nb["cat"], nb["ld"]
            else:
                src, ldsrc = nb["db_act"], c[l + 1]
            _ok(lib.rang_amd_nn_up(_p(src), ldsrc, 0, _p(b["cat"]), b["ld"], 0, B, H, W, c[l + 1], st))
            ka, kb = self.dec[l]
            ka.forward(b["cat"], b["ld"], 0, b["da_pre"], b["da_act"], c[l], 0, B, H, W)
            kb.forward(b["da_act"], c[l], 0, b["db_pre"], b["db_act"], c[l], 0, B, H, W)
        top = self.buf[0]["db_act"] if L > 1 else self.buf[0]["cat"]
        self.out.forward(top, c[0], 0, self.y, None, COUT, 0, B, self.H, self.W)
        return self.y

    # ------------------------------------------------------------ backward
    def backward(self, B):
        """Weight gradients (added to each conv's g) from the output gradient
        in self.g_y (FP32, channel 0)."""
        L, c, lib, st = self.levels, self.c, _lib(), _stream()
        H0, W0 = self.H, self.W

        def mask(pre):
            # the ReLU's gradient mask (none for the linear network)
            return None if self.linear else _p(pre)

        b0 = self.buf[0]
        top, ldtop = (b0["db_act"], c[0]) if L > 1 else (b0["cat"], c[0])
        gy16 = b0["gb16"]
        P0 = B * H0 * W0
        _ok(lib.rang_amd_nn_relu_bwd(_p(self.g_y), COUT, 0, None, 0, 0, _p(gy16), COUT, 0, P0, COUT, st))
        self.out.wgrad(top, ldtop, 0, gy16, COUT, 0, B, H0, W0)
        g_top = b0["g_d"] if L > 1 else b0["g_cat"]
        self.out.dgrad(gy16, COUT, 0, g_top, c[0], 0, B, H0, W0)
        # decoder, finest level first
        for l in range(0, L - 1):
            b, nb = self.buf[l], self.buf[l + 1]
            H, W = H0 >> l, W0 >> l
            P = B * H * W
            ka, kb = self.dec[l]
            g16 = b["gb16"]
            _ok(lib.rang_amd_nn_relu_bwd(_p(b["g_d"]), c[l], 0, mask(b["db_pre"]), c[l], 0, _p(g16), c[l], 0, P, c[l],
                                         st))
            kb.wgrad(b["da_act"], c[l], 0, g16, c[l], 0, B, H, W)
            kb.dgrad(g16, c[l], 0, b["g_a"], c[l], 0, B, H, W)
            _ok(lib.rang_amd_nn_relu_bwd(_p(b["g_a"]), c[l], 0, mask(b["da_pre"]), c[l], 0, _p(g16), c[l], 0, P, c[l],
                                         st))
            ka.wgrad(b["cat"], b["ld"], 0, g16, c[l], 0, B, H, W)
            ka.dgrad(g16, c[l], 0, b["g_cat"], b["ld"], 0, B, H, W)
            # the upsampled part: to the coarser level's output gradient
            g_dst = self.g_bottom if l + 1 == L - 1 else nb["g_d"]
            _ok(lib.rang_amd_nn_up_bwd(_p(b["g_cat"]), b["ld"], 0, _p(g_dst), c[l + 1], 0, B, H, W, c[l + 1], st))
        # encoder, coarsest level first
        for l in range(L - 1, -1, -1):
            b = self.buf[l]
            H, W = H0 >> l, W0 >> l
            P = B * H * W
            ka, kb = self.enc[l]
            g16 = b["gb16"]
            if l == L - 1:
                gsrc, ldg, offg = self.g_bottom, c[l], 0
                pre_ld, pre_off = b["ld"], 0
                if L == 1:
                    gsrc, ldg, offg = b["g_cat"], b["ld"], 0
            else:
                gsrc, ldg, offg = b["g_cat"], b["ld"], c[l + 1]
                pre_ld, pre_off = b["ld"], c[l + 1]
            _ok(lib.rang_amd_nn_relu_bwd(_p(gsrc), ldg, offg, mask(b["b_pre"]), pre_ld, pre_off, _p(g16), c[l], 0, P,
                                         c[l], st))
            kb.wgrad(b["a_act"], c[l], 0, g16, c[l], 0, B, H, W)
            kb.dgrad(g16, c[l], 0, b["g_a"], c[l], 0, B, H, W)
            _ok(lib.rang_amd_nn_relu_bwd(_p(b["g_a"]), c[l], 0, mask(b["a_pre"]), c[l], 0, _p(g16), c[l], 0, P, c[l],
                                         st))
            ka.wgrad(b["in"], b["in"].shape[1], 0, g16, c[l], 0, B, H, W)
            if l > 0:
                ka.dgrad(g16, c[l], 0, b["g_in"], c[l - 1], 0, B, H, W)
                pb = self.buf[l - 1]
                # the pooled skip of the finer level: add to its skip gradient
                _ok(lib.rang_amd_nn_pool_bwd(_p(b["g_in"]), c[l - 1], 0, _p(pb["g_cat"]), pb["ld"], c[l], B, H << 1,
                                             W << 1, c[l - 1], st))

    def zero_grad(self):
        for k in self.convs:
            k.g.zero_()

    def adam(self, lr):
        self.step += 1
        for k in self.convs:
            k.adam(self.step, lr)

    # --------------------------------------------------------------- state
    def state(self):
        return {f"w{i}": k.w.detach().cpu().numpy() for i, k in enumerate(self.convs)}

    def load(self, state):
        for i, k in enumerate(self.convs):
            k.w.copy_(torch.from_numpy(np.asarray(state[f"w{i}"], dtype=np.float32)))
            k.refresh()


def _psf_image(measurement):
    """The plan's PSF (vis2dirty of unit visibilities), peak one."""
    plan = measurement.plan
    ones = torch.ones_like(measurement.vis)
    psf = plan.vis2dirty(ones)
    return psf / psf.abs().max()


class Preconditioner:
    """M(r) = |r| (N(r / |r|) - N(-r / |r|)) / 2 with the network N and the
    plan's PSF as its second input channel."""

    def __init__(self, measurement, net):
        self.m, self.net = measurement, net
        plan = measurement.plan
        self.real = plan.real
        side = measurement.side
        if (net.H, net.W) != (side, side) or net.batch < 2:
            raise ValueError("the network's buffers must hold two images of the measurement's size")
        self.hw = side * side
        psf = _psf_image(measurement).contiguous()
        self.psf = psf
        inp = net.buf[0]["in"]
        inp.zero_()
        fn = _lib().rang_amd_nn_pack_f64 if self.real == torch.float64 else _lib().rang_amd_nn_pack_f32
        # the PSF channel of both images
        two = torch.stack([psf, psf]).contiguous()
        _ok(fn(_p(two), None, _p(inp), CIN, 1, 2, self.hw, _stream()))
        self.scale = torch.empty(2, dtype=torch.float64, device=psf.device)
        self.norm = torch.empty(1, dtype=torch.float64, device=psf.device)
        self.pair = torch.empty((2, side, side), dtype=self.real, device=psf.device)
        self.applies = 0

    def __call__(self, r, out=None):
        lib, st = _lib(), _stream()
        f64 = self.real == torch.float64
        dot = lib.rang_amd_nn_image_dot_f64 if f64 else lib.rang_amd_nn_image_dot_f32
        _ok(dot(_p(r), None, _p(self.norm), 1, self.hw, st))
        nrm = math.sqrt(float(self.norm[0]) / self.hw)
        z = out if out is not None else torch.empty_like(r)
        if not nrm > 0.0:
            z.zero_()
            return z
        self.scale[0] = 1.0 / nrm
        self.scale[1] = -1.0 / nrm
        self.pair[0].copy_(r)
        self.pair[1].copy_(r)
        pack = lib.rang_amd_nn_pack_f64 if f64 else lib.rang_amd_nn_pack_f32
        _ok(pack(_p(self.pair), _p(self.scale), _p(self.net.buf[0]["in"]), CIN, 0, 2, self.hw, st))
        y = self.net.forward(2)
        unpack = lib.rang_amd_nn_unpack_f64 if f64 else lib.rang_amd_nn_unpack_f32
        self.scale[0] = 0.5 * nrm
        self.scale[1] = -0.5 * nrm
        _ok(unpack(_p(y), COUT, 0, _p(self.scale), _p(self.pair), 2, self.hw, st))
        axpby = lib.rang_amd_nn_axpby_f64 if f64 else lib.rang_amd_nn_axpby_f32
        z.copy_(self.pair[0])
        _ok(axpby(_p(self.pair[1]), _p(z), 1.0, 1.0, self.hw, st))
        self.applies += 1
        return z


class Trainer:
    """Trains the network on residuals the gridder generates: random images x
    (sums of Gaussians from point-like to smooth, of either sign, and positive
    point-source skies) and their Gram images r = G x from the prepared plan,
    each scaled to unit RMS. With loss="residual" (the default) the loss is
    |G N(r) - r|^2 / (2 |r|^2), the preconditioner's own residual, whose
    gradient G (G N(r) - r) / |r|^2 takes two more Gram applications per
    image; with loss="image" it is |N(r) - x|^2 / (2 |x|^2)."""

    def __init__(self, measurement, net, *, seed=7, components=(64, 24), widths=(0.5, 24.0), loss="residual"):
        if loss not in ("residual", "image"):
            raise ValueError("loss is 'residual' or 'image'")
        self.kind = loss
        self.m, self.net = measurement, net
        self.side = measurement.side
        self.hw = self.side * self.side
        self.seed = seed
        self.components = components
        self.widths = widths
        dev = net.y.device
        B = net.batch
        self.B = B
        self.x = torch.empty((B, self.side, self.side), dtype=torch.float64, device=dev)
        self.gx = torch.empty_like(self.x)
        self.z = torch.empty_like(self.x)
        self.gz = torch.empty_like(self.x)
        self.gu = torch.empty_like(self.x)
        self.t = torch.empty((B, self.side, self.side), dtype=torch.float32, device=dev)
        self.s = torch.empty(B, dtype=torch.float64, device=dev)
        self.inv = torch.empty(B, dtype=torch.float64, device=dev)
        self.loss = torch.empty(1, dtype=torch.float64, device=dev)
        psf = _psf_image(measurement).to(torch.float64).contiguous()
        inp = net.buf[0]["in"]
        inp.zero_()
        allpsf = psf.expand(B, -1, -1).contiguous()
        _ok(_lib().rang_amd_nn_pack_f64(_p(allpsf), None, _p(inp), CIN, 1, B, self.hw, _stream()))
        self.batches = 0
        self.gram_applies = 0

    def sample(self):
        """A batch of pairs: half signed multi-scale fields, half positive skies."""
        lib, st = _lib(), _stream()
        B, side = self.B, self.side
        half = B // 2
        k_s, k_p = self.components
        s0 = (self.seed * 1000003 + self.batches) & 0xFFFFFFFFFFFFFFFF
        _ok(lib.rang_amd_nn_sky_f64(_p(self.x), half, side, side, k_s, self.widths[0], self.widths[1], 0, s0, st))
        if B - half > 0:
            _ok(lib.rang_amd_nn_sky_f64(_p(self.x[half:]), B - half, side, side, k_p, 0.5, 1.5, 1,
                                        s0 ^ 0x5bd1e995, st))
        gram = self.m.plan.gram
        for n in range(B):
            xin = self.x[n] if self.m.plan.real == torch.float64 else self.x[n].to(self.m.plan.real)
            g = gram(xin.contiguous())
            self.gx[n].copy_(g)
        self.gram_applies += B
        self.batches += 1

    def _gram_each(self, src, dst):
        gram = self.m.plan.gram
        real = self.m.plan.real
        for n in range(self.B):
            xin = src[n] if real == torch.float64 else src[n].to(real)
            dst[n].copy_(gram(xin.contiguous()))
        self.gram_applies += self.B

    def step(self, lr):
        if self.kind == "residual":
            return self._step_residual(lr)
        lib, st = _lib(), _stream()
        B, hw = self.B, self.hw
        self.sample()
        # per image: s = 1 / rms(G x)
        _ok(lib.rang_amd_nn_image_dot_f64(_p(self.gx), None, _p(self.s), B, hw, st))
        ss = self.s.cpu().numpy()
        rms = np.sqrt(ss / hw)
        self.s.copy_(torch.from_numpy(np.where(rms > 0, 1.0 / np.where(rms > 0, rms, 1.0), 0.0)))
        _ok(lib.rang_amd_nn_pack_f64(_p(self.gx), _p(self.s), _p(self.net.buf[0]["in"]), CIN, 0, B, hw, st))
        # targets x s, loss weights 1 / (B |x s|^2)
        _ok(lib.rang_amd_nn_scale_images(_p(self.x), _p(self.s), _p(self.t), B, hw, st))
        _ok(lib.rang_amd_nn_image_dot_f32(_p(self.t), None, _p(self.inv), B, hw, st))
        tt = self.inv.cpu().numpy()
        self.inv.copy_(torch.from_numpy(np.where(tt > 0, 1.0 / (np.where(tt > 0, tt, 1.0) * B), 0.0)))
        y = self.net.forward(B)
        _ok(lib.rang_amd_nn_loss(_p(y), COUT, 0, _p(self.t), _p(self.inv), _p(self.net.g_y), COUT, B, hw,
                                 _p(self.loss), st))
        self.net.zero_grad()
        self.net.backward(B)
        self.net.adam(lr)
        return float(self.loss[0])

    def collect(self, rhs_count=8, iterations=200, keep=(0, 2, 5, 10, 20, 40, 80, 120, 160, 199)):
        """A pool of conjugate-gradient residuals (plain CG on the Gram
        operator from zero, right-hand sides G x of random images), kept at the
        iterations in `keep`: the inputs a preconditioner meets inside the
        solve. One Gram application per CG iteration."""
        lib, st = _lib(), _stream()
        plan = self.m.plan
        hw = self.hw
        pool = []
        acc = torch.empty(1, dtype=torch.float64, device=self.x.device)

        def dot(a, b):
            _ok(lib.rang_amd_nn_image_dot_f64(_p(a), _p(b), _p(acc), 1, hw, st))
            return float(acc[0])

        for j in range(0, rhs_count, self.B):
            self.sample()
            for n in range(min(self.B, rhs_count - j)):
                r = self.gx[n].clone()
                p = r.clone()
                x = torch.zeros_like(r)
                rr = dot(r, r)
                for k in range(iterations):
                    if k in keep:
                        pool.append((r / math.sqrt(rr / hw)).contiguous())
                    q = plan.gram(p)
                    self.gram_applies += 1
                    pq = dot(p, q)
                    if not pq > 0:
                        break
                    a = rr / pq
                    _ok(lib.rang_amd_nn_axpby_f64(_p(p), _p(x), a, 1.0, hw, st))
                    _ok(lib.rang_amd_nn_axpby_f64(_p(q), _p(r), -a, 1.0, hw, st))
                    rn = dot(r, r)
                    _ok(lib.rang_amd_nn_axpby_f64(_p(r), _p(p), 1.0, rn / rr, hw, st))
                    rr = rn
        self.pool = pool
        return len(pool)

    def _step_residual(self, lr):
        lib, st = _lib(), _stream()
        B, hw = self.B, self.hw
        self.sample()
        pool = getattr(self, "pool", None)
        if pool:
            # half the batch from the pool of CG residuals
            rng = np.random.default_rng(self.seed * 7919 + self.batches)
            for n in range(B // 2, B):
                self.gx[n].copy_(pool[int(rng.integers(len(pool)))])
        _ok(lib.rang_amd_nn_image_dot_f64(_p(self.gx), None, _p(self.s), B, hw, st))
        ss = self.s.cpu().numpy()
        rms = np.sqrt(ss / hw)
        scale = np.where(rms > 0, 1.0 / np.where(rms > 0, rms, 1.0), 0.0)
        self.s.copy_(torch.from_numpy(scale))
        _ok(lib.rang_amd_nn_pack_f64(_p(self.gx), _p(self.s), _p(self.net.buf[0]["in"]), CIN, 0, B, hw, st))
        y = self.net.forward(B)
        _ok(lib.rang_amd_nn_unpack_f64(_p(y), COUT, 0, None, _p(self.z), B, hw, st))
        # u = G z - s r per image; loss = sum |u|^2 / (2 hw B) (|s r|^2 = hw)
        self._gram_each(self.z, self.gz)
        for n in range(B):
            _ok(lib.rang_amd_nn_axpby_f64(_p(self.gx[n]), _p(self.gz[n]), -float(scale[n]), 1.0, hw, st))
        _ok(lib.rang_amd_nn_image_dot_f64(_p(self.gz), None, _p(self.inv), B, hw, st))
        loss = float(self.inv.sum().cpu()) / (2.0 * hw * B)
        # dL/dz = G u / (hw B)
        self._gram_each(self.gz, self.gu)
        self.inv.fill_(1.0 / (hw * B))
        _ok(lib.rang_amd_nn_grad_images(_p(self.gu), _p(self.inv), _p(self.net.g_y), COUT, 0, B, hw, st))
        self.net.zero_grad()
        self.net.backward(B)
        self.net.adam(lr)
        return loss

    def train(self, steps, lr=1e-3, lr_end=1e-4, log=None):
        losses = []
        for k in range(steps):
            a = lr * (lr_end / lr) ** (k / max(steps - 1, 1))
            losses.append(self.step(a))
            if log and (k % 25 == 0 or k == steps - 1):
                log(k, losses[-1])
        return losses


def preconditioned_cg(measurement, precond, rhs, iterations, tolerance, *, restart=0, stats=None, truncation=1):
    """Flexible conjugate gradients (Notay 2000; the Polak-Ribiere form of
    Golub and Ye 1999) on A^H W A x = rhs from zero with z = precond(r): at
    most `iterations` Gram applications, stopping at |r| <= tolerance |rhs|
    on the recursive residual. Where precond(r) is not a descent direction
    (<r, z> <= 1e-3 |r| |z|) the step takes z = r, the plain CG direction;
    `restart` > 0 restarts the directions every so many steps; with
    `truncation` m > 1 each direction is made G-conjugate to the last m (the
    truncated FCG(m) of Notay 2000) instead of the one-term recurrence. Returns (x,
    Gram applications, the true relative residual |rhs - A^H W A x| / |rhs|
    from one more exact application); `stats` (a dict) receives the number
    of fallback steps."""
    plan = measurement.plan
    real = plan.real
    lib, st = _lib(), _stream()
    f64 = real == torch.float64
    axpby = lib.rang_amd_nn_axpby_f64 if f64 else lib.rang_amd_nn_axpby_f32
    idot = lib.rang_amd_nn_image_dot_f64 if f64 else lib.rang_amd_nn_image_dot_f32
    n = rhs.numel()
    acc = torch.empty(1, dtype=torch.float64, device=rhs.device)

    def dot(a, b):
        _ok(idot(_p(a), _p(b), _p(acc), 1, n, st))
        return float(acc[0])

    b = rhs.to(real).contiguous()
    x = torch.zeros_like(b)
    r = b.clone()
    nb = math.sqrt(dot(b, b))
    if nb == 0.0:
        return x, 0, 0.0
    p = torch.zeros_like(b)
    r_prev = torch.zeros_like(b)
    rz_prev = 1.0
    hist = []  # (p_j, q_j = G p_j, <p_j, q_j>) of the last `truncation` steps
    used = 0
    fallbacks = 0
    k = 0
    rr = nb * nb
    while used < iterations and math.sqrt(rr) > tolerance * nb:
        z = precond(r)
        rz = dot(r, z)
        zz = dot(z, z)
        if not rz > 1e-3 * math.sqrt(rr * zz):
            z = r.clone()
            rz = rr
            fallbacks += 1
        if k == 0 or (restart and k % restart == 0):
            p.copy_(z)
            hist.clear()
        elif truncation > 1:
            p = z.clone()
            for pj, qj, pqj in hist:
                _ok(axpby(_p(pj), _p(p), -dot(z, qj) / pqj, 1.0, n, st))
        else:
            beta = (rz - dot(z, r_prev)) / rz_prev
            _ok(axpby(_p(z), _p(p), 1.0, beta, n, st))  # p = z + beta p
        q = plan.gram(p)
        used += 1
        pq = dot(p, q)
        if not pq > 0.0:
            break
        alpha = dot(r, p) / pq if truncation > 1 else rz / pq
        if truncation > 1:
            hist.append((p, q, pq))
            if len(hist) > truncation:
                hist.pop(0)
        _ok(axpby(_p(p), _p(x), alpha, 1.0, n, st))
        r_prev.copy_(r)
        _ok(axpby(_p(q), _p(r), -alpha, 1.0, n, st))
        rz_prev = rz
        rr = dot(r, r)
        k += 1
    true_r = plan.gram(x)
    _ok(axpby(_p(b), _p(true_r), 1.0, -1.0, n, st))  # rhs - G x
    if stats is not None:
        stats["fallbacks"] = fallbacks
    return x, used, math.sqrt(dot(true_r, true_r)) / nb


# --------------------------------------------------------------------------
# A global path for the network: the circulant inverse filter of the plan's
# PSF (rang.gram.GramOperator.inverse_filter, rang-core's GramOperator::
# inverse_filter, core/src/gridder_gram.rs:488-509) as the preconditioner's
# linear part, and the U-Net as a learned correction to it.


def _rms(lib, img, acc, hw, f64):
    dot = lib.rang_amd_nn_image_dot_f64 if f64 else lib.rang_amd_nn_image_dot_f32
    _ok(dot(_p(img), None, _p(acc), 1, hw, _stream()))
    return math.sqrt(float(acc[0]) / hw)


class FilteredPreconditioner:
    """z = M r + c (N(u) - N(-u)) / 2 with M the inverse filter (symmetric
    positive definite, global), N the network on u = (r / rms(r), PSF,
    M r / c) and c = rms(M r). Odd and homogeneous of degree one in r;
    with an untrained or absent network it is M."""

    def __init__(self, measurement, net, filt):
        self.m, self.net, self.filt = measurement, net, filt
        plan = measurement.plan
        self.real = plan.real
        side = measurement.side
        if net is not None and ((net.H, net.W) != (side, side) or net.batch < 2):
            raise ValueError("the network's buffers must hold two images of the measurement's size")
        self.hw = side * side
        dev = plan.device
        self.acc = torch.empty(1, dtype=torch.float64, device=dev)
        self.scale = torch.empty(2, dtype=torch.float64, device=dev)
        self.pair = torch.empty((2, side, side), dtype=self.real, device=dev)
        self.mr = torch.empty((side, side), dtype=self.real, device=dev)
        self.applies = 0
        if net is not None:
            inp = net.buf[0]["in"]
            inp.zero_()
            psf = _psf_image(measurement).to(self.real).contiguous()
            two = torch.stack([psf, psf]).contiguous()
            pack = _lib().rang_amd_nn_pack_f64 if self.real == torch.float64 else _lib().rang_amd_nn_pack_f32
            _ok(pack(_p(two), None, _p(inp), CIN, 1, 2, self.hw, _stream()))

    def __call__(self, r, out=None):
        lib, st = _lib(), _stream()
        f64 = self.real == torch.float64
        z = out if out is not None else torch.empty_like(r)
        self.filt.apply(r, out=self.mr)
        self.applies += 1
        if self.net is None:
            z.copy_(self.mr)
            return z
        nr = _rms(lib, r, self.acc, self.hw, f64)
        c = _rms(lib, self.mr, self.acc, self.hw, f64)
        if not (nr > 0.0 and c > 0.0):
            z.copy_(self.mr)
            return z
        pack = lib.rang_amd_nn_pack_f64 if f64 else lib.rang_amd_nn_pack_f32
        inp = self.net.buf[0]["in"]
        self.pair[0].copy_(r)
        self.pair[1].copy_(r)
        self.scale[0], self.scale[1] = 1.0 / nr, -1.0 / nr
        _ok(pack(_p(self.pair), _p(self.scale), _p(inp), CIN, 0, 2, self.hw, st))
        self.pair[0].copy_(self.mr)
        self.pair[1].copy_(self.mr)
        self.scale[0], self.scale[1] = 1.0 / c, -1.0 / c
        _ok(pack(_p(self.pair), _p(self.scale), _p(inp), CIN, 2, 2, self.hw, st))
        y = self.net.forward(2)
        unpack = lib.rang_amd_nn_unpack_f64 if f64 else lib.rang_amd_nn_unpack_f32
        self.scale[0], self.scale[1] = 0.5 * c, -0.5 * c
        _ok(unpack(_p(y), COUT, 0, _p(self.scale), _p(self.pair), 2, self.hw, st))
        axpby = lib.rang_amd_nn_axpby_f64 if f64 else lib.rang_amd_nn_axpby_f32
        z.copy_(self.mr)
        _ok(axpby(_p(self.pair[0]), _p(z), 1.0, 1.0, self.hw, st))
        _ok(axpby(_p(self.pair[1]), _p(z), 1.0, 1.0, self.hw, st))
        return z


class FilteredTrainer:
    """Trains the network as the correction of FilteredPreconditioner: for
    residual images r (the operator's images G x of random skies, and
    residuals of the filter-preconditioned CG on random right-hand sides),
    minimise |G (M r^ + c y) - r^|^2 / (2 |r^|^2), r^ = r / rms(r), c =
    rms(M r^), y the network's output on (r^, PSF, M r^ / c); G is `gram`,
    an operator with apply(x, out) (the exact normal operator of the plan, or
    the normal operator without visibilities, rang.gram.GramOperator, which
    costs no gridding). Counts G's applications (gram_applies) and the exact
    operator's (exact_applies, those of collect)."""

    def __init__(self, measurement, net, filt, gram, *, seed=7, components=(64, 24), widths=(0.5, 24.0)):
        self.m, self.net, self.filt, self.G = measurement, net, filt, gram
        self.side = measurement.side
        self.hw = self.side * self.side
        self.seed, self.components, self.widths = seed, components, widths
        dev = net.y.device
        B = net.batch
        self.B = B
        real = measurement.plan.real
        self.real = real
        self.x = torch.empty((B, self.side, self.side), dtype=torch.float64, device=dev)
        self.r = torch.empty((B, self.side, self.side), dtype=real, device=dev)
        self.mr = torch.empty_like(self.r)
        self.y = torch.empty_like(self.r)
        self.u = torch.empty_like(self.r)
        self.gu = torch.empty_like(self.r)
        self.t = torch.empty((B, self.side, self.side), dtype=torch.float64, device=dev)
        self.s = torch.empty(B, dtype=torch.float64, device=dev)
        self.acc = torch.empty(B, dtype=torch.float64, device=dev)
        psf = _psf_image(measurement).to(real).contiguous()
        inp = net.buf[0]["in"]
        inp.zero_()
        allpsf = psf.expand(B, -1, -1).contiguous()
        pack = _lib().rang_amd_nn_pack_f64 if real == torch.float64 else _lib().rang_amd_nn_pack_f32
        _ok(pack(_p(allpsf), None, _p(inp), CIN, 1, B, self.hw, _stream()))
        self.batches = 0
        self.gram_applies = 0
        self.exact_applies = 0
        self.pool = []

    def _g(self, src, dst):
        for n in range(self.B):
            self.G.apply(src[n].contiguous(), out=dst[n])
        self.gram_applies += self.B

    def collect(self, rhs_count=8, iterations=60, keep=(0, 1, 2, 4, 8, 16, 32, 59)):
        """Residuals of CG preconditioned by the filter alone, on right-hand
        sides G x of random skies (G the exact normal operator here, as in the
        solve), kept at the iterations in `keep`."""
        plan = self.m.plan
        lib, st = _lib(), _stream()
        f64 = self.real == torch.float64
        idot = lib.rang_amd_nn_image_dot_f64 if f64 else lib.rang_amd_nn_image_dot_f32
        axpby = lib.rang_amd_nn_axpby_f64 if f64 else lib.rang_amd_nn_axpby_f32
        acc = torch.empty(1, dtype=torch.float64, device=self.x.device)
        hw = self.hw

        def dot(a, b):
            _ok(idot(_p(a), _p(b), _p(acc), 1, hw, st))
            return float(acc[0])

        for j in range(rhs_count):
            self._sky(1)
            r = plan.gram(self.x[0].to(self.real).contiguous()).clone()
            self.exact_applies += 1
            x = torch.zeros_like(r)
            z = self.filt.apply(r)
            p = z.clone()
            rz = dot(r, z)
            for k in range(iterations):
                if k in keep:
                    rr = dot(r, r)
                    self.pool.append((r * (1.0 / math.sqrt(rr / hw))).contiguous())
                q = plan.gram(p)
                self.exact_applies += 1
                pq = dot(p, q)
                if not pq > 0:
                    break
                a = rz / pq
                _ok(axpby(_p(p), _p(x), a, 1.0, hw, st))
                _ok(axpby(_p(q), _p(r), -a, 1.0, hw, st))
                z = self.filt.apply(r)
                rz_n = dot(r, z)
                _ok(axpby(_p(z), _p(p), 1.0, rz_n / rz, hw, st))
                rz = rz_n
        return len(self.pool)

    def _sky(self, count):
        lib, st = _lib(), _stream()
        s0 = (self.seed * 1000003 + self.batches) & 0xFFFFFFFFFFFFFFFF
        self.batches += 1
        half = count // 2 if count > 1 else 0
        if half:
            _ok(lib.rang_amd_nn_sky_f64(_p(self.x), half, self.side, self.side, self.components[0], self.widths[0],
                                        self.widths[1], 0, s0, st))
        _ok(lib.rang_amd_nn_sky_f64(_p(self.x[half:]), count - half, self.side, self.side, self.components[1], 0.5,
                                    1.5, 1, s0 ^ 0x5bd1e995, st))

    def step(self, lr):
        lib, st = _lib(), _stream()
        B, hw = self.B, self.hw
        f64 = self.real == torch.float64
        idot = lib.rang_amd_nn_image_dot_f64 if f64 else lib.rang_amd_nn_image_dot_f32
        axpby = lib.rang_amd_nn_axpby_f64 if f64 else lib.rang_amd_nn_axpby_f32
        pack = lib.rang_amd_nn_pack_f64 if f64 else lib.rang_amd_nn_pack_f32
        unpack = lib.rang_amd_nn_unpack_f64 if f64 else lib.rang_amd_nn_unpack_f32
        # half the batch G x of random skies, half from the pool
        self._sky(B)
        self._g(self.x.to(self.real), self.r)
        if self.pool:
            rng = np.random.default_rng(self.seed * 7919 + self.batches)
            for n in range(B // 2, B):
                self.r[n].copy_(self.pool[int(rng.integers(len(self.pool)))])
        _ok(idot(_p(self.r), None, _p(self.acc), B, hw, st))
        nr = np.sqrt(self.acc.cpu().numpy() / hw)
        inv_r = np.where(nr > 0, 1.0 / np.where(nr > 0, nr, 1.0), 0.0)
        for n in range(B):
            _ok(axpby(_p(self.r[n]), _p(self.r[n]), float(inv_r[n]) - 1.0, 1.0, hw, st))  # r^ = r / rms(r)
            self.filt.apply(self.r[n], out=self.mr[n])
        _ok(idot(_p(self.mr), None, _p(self.acc), B, hw, st))
        c = np.sqrt(self.acc.cpu().numpy() / hw)
        inv_c = np.where(c > 0, 1.0 / np.where(c > 0, c, 1.0), 0.0)
        inp = self.net.buf[0]["in"]
        _ok(pack(_p(self.r), None, _p(inp), CIN, 0, B, hw, st))
        self.s.copy_(torch.from_numpy(inv_c))
        _ok(pack(_p(self.mr), _p(self.s), _p(inp), CIN, 2, B, hw, st))
        yout = self.net.forward(B)
        self.s.copy_(torch.from_numpy(c))
        _ok(unpack(_p(yout), COUT, 0, _p(self.s), _p(self.y), B, hw, st))
        # z = M r^ + c y; u = G z - r^
        _ok(axpby(_p(self.mr), _p(self.y), 1.0, 1.0, B * hw, st))
        self._g(self.y, self.u)
        _ok(axpby(_p(self.r), _p(self.u), -1.0, 1.0, B * hw, st))
        _ok(idot(_p(self.u), None, _p(self.acc), B, hw, st))
        loss = float(self.acc.sum().cpu()) / (2.0 * hw * B)
        # dL/dy = c G u / (hw B)
        self._g(self.u, self.gu)
        self.t.copy_(self.gu)
        self.s.copy_(torch.from_numpy(c / (hw * B)))
        _ok(lib.rang_amd_nn_grad_images(_p(self.t), _p(self.s), _p(self.net.g_y), COUT, 0, B, hw, st))
        self.net.zero_grad()
        self.net.backward(B)
        self.net.adam(lr)
        return loss

    def train(self, steps, lr=1e-3, lr_end=1e-4, log=None):
        losses = []
        for k in range(steps):
            a = lr * (lr_end / lr) ** (k / max(steps - 1, 1))
            losses.append(self.step(a))
            if log and (k % 25 == 0 or k == steps - 1):
                log(k, losses[-1])
        return losses


def safeguarded_fcg(measurement, precond, fallback, rhs, iterations, tolerance, *, window=8, factor=0.5, stats=None):
    """Flexible CG (Notay 2000, one-term) with `precond`, judged by the exact
    normal operator: when over `window` steps the residual falls by less
    than `factor`, or a direction is not one of descent, the rest of the
    solve is conjugate gradients preconditioned by `fallback` (symmetric
    positive definite; None: plain CG) from the current iterate. Returns
    (x, Gram applications, |rhs - A^H W A x| / |rhs| from one more exact
    application); `stats` gets the step at which the fallback took over."""
    plan = measurement.plan
    real = plan.real
    lib, st = _lib(), _stream()
    f64 = real == torch.float64
    axpby = lib.rang_amd_nn_axpby_f64 if f64 else lib.rang_amd_nn_axpby_f32
    idot = lib.rang_amd_nn_image_dot_f64 if f64 else lib.rang_amd_nn_image_dot_f32
    n = rhs.numel()
    acc = torch.empty(1, dtype=torch.float64, device=rhs.device)

    def dot(a, b):
        _ok(idot(_p(a), _p(b), _p(acc), 1, n, st))
        return float(acc[0])

    b = rhs.to(real).contiguous()
    x = torch.zeros_like(b)
    r = b.clone()
    nb = math.sqrt(dot(b, b))
    if nb == 0.0:
        return x, 0, 0.0
    p = torch.zeros_like(b)
    r_prev = torch.zeros_like(b)
    rz_prev = 1.0
    used, k = 0, 0
    rr = nb * nb
    hist = [math.sqrt(rr)]
    switched_at = None
    pre = precond
    while used < iterations and math.sqrt(rr) > tolerance * nb:
        if pre is None:
            z = r.clone()
        else:
            z = pre(r)
        rz = dot(r, z)
        if not rz > 0.0 and switched_at is None:
            switched_at, pre, k = used, fallback, 0
            continue
        if k == 0:
            p.copy_(z)
        else:
            beta = (rz - dot(z, r_prev)) / rz_prev if switched_at is None else rz / rz_prev
            _ok(axpby(_p(z), _p(p), 1.0, beta, n, st))
        q = plan.gram(p)
        used += 1
        pq = dot(p, q)
        if not pq > 0.0:
            break
        alpha = rz / pq
        _ok(axpby(_p(p), _p(x), alpha, 1.0, n, st))
        r_prev.copy_(r)
        _ok(axpby(_p(q), _p(r), -alpha, 1.0, n, st))
        rz_prev = rz
        rr = dot(r, r)
        hist.append(math.sqrt(rr))
        k += 1
        if switched_at is None and len(hist) > window and hist[-1] > factor * hist[-1 - window]:
            switched_at, pre, k = used, fallback, 0
    true_r = plan.gram(x)
    _ok(axpby(_p(b), _p(true_r), 1.0, -1.0, n, st))
    if stats is not None:
        stats["switched_at"] = switched_at
    return x, used, math.sqrt(dot(true_r, true_r)) / nb


class SurrogateSolve:
    """The preconditioner of the inner-outer solve: z ~ G~^{-1} r by conjugate
    gradients on the surrogate G~ (the w-aware normal operator without
    visibilities, rang.gram.GramOperator; rang-core's GramOperator::new,
    core/src/gridder_gram.rs:226-407) preconditioned by `filt` (its circulant
    inverse filter, :488-509), stopped at relative residual `tau` or after
    `inner_max` steps. With `learned` (a FilteredPreconditioner) the inner
    solve starts from the network's estimate instead of zero. Counts the
    surrogate's and the network's applications."""

    def __init__(self, surrogate, filt, *, tau=0.1, inner_max=400, learned=None):
        self.G, self.M, self.tau, self.inner_max, self.learned = surrogate, filt, tau, inner_max, learned
        self.surrogate_applies = 0
        self.learned_applies = 0
        self.inner_steps = []

    def __call__(self, r, out=None):
        lib, st = _lib(), _stream()
        f64 = r.dtype == torch.float64
        axpby = lib.rang_amd_nn_axpby_f64 if f64 else lib.rang_amd_nn_axpby_f32
        idot = lib.rang_amd_nn_image_dot_f64 if f64 else lib.rang_amd_nn_image_dot_f32
        n = r.numel()
        acc = torch.empty(1, dtype=torch.float64, device=r.device)

        def dot(a, b):
            _ok(idot(_p(a), _p(b), _p(acc), 1, n, st))
            return float(acc[0])

        z = out if out is not None else torch.empty_like(r)
        res = r.clone()
        if self.learned is not None:
            self.learned(r, out=z)
            self.learned_applies += 1
            q = self.G.apply(z)
            self.surrogate_applies += 1
            _ok(axpby(_p(q), _p(res), -1.0, 1.0, n, st))
        else:
            z.zero_()
        n0 = math.sqrt(dot(r, r))
        s = self.M.apply(res)
        p = s.clone()
        rs = dot(res, s)
        k = 0
        while k < self.inner_max and math.sqrt(dot(res, res)) > self.tau * n0:
            q = self.G.apply(p)
            self.surrogate_applies += 1
            k += 1
            pq = dot(p, q)
            if not pq > 0.0:
                break
            a = rs / pq
            _ok(axpby(_p(p), _p(z), a, 1.0, n, st))
            _ok(axpby(_p(q), _p(res), -a, 1.0, n, st))
            s = self.M.apply(res)
            rs_n = dot(res, s)
            _ok(axpby(_p(s), _p(p), 1.0, rs_n / rs, n, st))
            rs = rs_n
        self.inner_steps.append(k)
        return z


def inner_outer_solve(measurement, surrogate, filt, rhs, iterations, tolerance, *, tau=0.1, inner_max=400,
                      learned=None, window=4, factor=0.5):
    """Flexible CG on the exact normal operator A^H W A (every residual
    judged by it), each step preconditioned by SurrogateSolve; when over
    `window` outer steps the residual falls by less than `factor`, the rest
    of the solve is plain CG (core/src/clean.rs:263-306), the fallback.
    Returns (x, counts, rel): counts has the exact operator's applications
    in the iteration ("exact"; one more forms the judged residual rel), the
    surrogate's and the network's, the inner steps per outer step and the
    step at which the fallback took over (None if it did not)."""
    pre = SurrogateSolve(surrogate, filt, tau=tau, inner_max=inner_max, learned=learned)
    st = {}
    x, used, rel = safeguarded_fcg(measurement, pre, None, rhs, iterations, tolerance, window=window,
                                   factor=factor, stats=st)
    counts = {"exact": used, "surrogate": pre.surrogate_applies, "learned": pre.learned_applies,
              "inner_steps": pre.inner_steps, "switched_at": st.get("switched_at")}
    return x, counts, rel
