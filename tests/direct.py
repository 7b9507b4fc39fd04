"""FP64 direct transforms on the GPU (torch) for the parity tests at the
workloads' sizes, and the error measures the tests and the benchmark report.

The closed forms are the gridder's conventions (core/src/wgridder.rs:30-36):
dirty(l, m) = sum_k w_k Re[V_k exp(+2 pi i (u l + v m + w (n - 1)) f/c)],
V_k = w_k sum_p I_p exp(-2 pi i (u l_p + v m_p + w (n_p - 1)) f/c), with
l = (i - nx/2) dx, m = (j - ny/2) dy."""

import math

import numpy as np

C_LIGHT = 299792458.0


def errors(got, ref):
    """(L2, Linf) of got - ref relative to ref: |got - ref|_2 / |ref|_2 and
    max |got - ref| / max |ref|, in float64."""
    g = np.asarray(got, dtype=np.complex128 if np.iscomplexobj(got) else np.float64).ravel()
    r = np.asarray(ref, dtype=g.dtype).ravel()
    d = g - r
    return (float(np.linalg.norm(d) / np.linalg.norm(r)), float(np.abs(d).max() / np.abs(r).max()))


def lm(i, j, nx, ny, dx, dy):
    import torch

    l = (i.to(torch.float64) - nx // 2) * dx
    m = (j.to(torch.float64) - ny // 2) * dy
    return l, m


def dirty_pixels(uvw, freq, vis, wgt, i, j, nx, ny, dx, dy, chunk=4096):
    """The dirty image at pixels (i[k], j[k]) (torch tensors on the GPU) by
    the FP64 direct sum over every visibility; uvw (nrow, 3), freq (nchan),
    vis and wgt (nrow, nchan) torch tensors (wgt may be None)."""
    import torch

    l, m = lm(i, j, nx, ny, dx, dy)
    nm1 = torch.sqrt(1.0 - l * l - m * m) - 1.0
    foc = freq.to(torch.float64) / C_LIGHT
    out = torch.zeros(i.numel(), dtype=torch.float64, device=i.device)
    for r0 in range(0, uvw.shape[0], chunk):
        u = uvw[r0:r0 + chunk].to(torch.float64)
        v = vis[r0:r0 + chunk].to(torch.complex128)
        if wgt is not None:
            v = v * wgt[r0:r0 + chunk].to(torch.float64)
        path = u[:, 0:1] * l[None, :] + u[:, 1:2] * m[None, :] + u[:, 2:3] * nm1[None, :]
        ph = 2.0 * math.pi * path[:, None, :] * foc[None, :, None]
        out += (v[:, :, None].real * torch.cos(ph) - v[:, :, None].imag * torch.sin(ph)).sum(dim=(0, 1))
    return out


def vis_of_points(uvw, freq, rows, chans, pi, pj, flux, nx, ny, dx, dy):
    """Visibilities (complex128) at (rows[k], chans[k]) of an image that is
    `flux` at pixels (pi, pj) and zero elsewhere, by the FP64 direct sum."""
    import torch

    l, m = lm(pi, pj, nx, ny, dx, dy)
    nm1 = torch.sqrt(1.0 - l * l - m * m) - 1.0
    u = uvw[rows].to(torch.float64)
    foc = (freq.to(torch.float64) / C_LIGHT)[chans]
    path = u[:, 0:1] * l[None, :] + u[:, 1:2] * m[None, :] + u[:, 2:3] * nm1[None, :]
    ph = -2.0 * math.pi * path * foc[:, None]
    f = flux.to(torch.float64)[None, :]
    return torch.complex((f * torch.cos(ph)).sum(1), (f * torch.sin(ph)).sum(1))
