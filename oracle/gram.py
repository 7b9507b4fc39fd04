"""ORACLE — TEST INFRASTRUCTURE ONLY.

Closed forms for the normal operator without visibilities (rang-core
core/src/gridder_gram.rs), used by tests/ to check the GPU operator; the
product library never imports this module.

* gram_direct: the Gram matrix applied by its direct sum in the gridder's
  conventions (the `direct` of gridder_gram.rs:781-814):
  (G m)(l) = sum_p W_p sum_l' m(l') cos 2 pi (u_p (l - l') + v_p (m - m') +
  w_p (x(l) - x(l'))), x = n - 1, with v and w negated by flip_v / flip_w.
* expansion_error: |exp(i theta xi s) - sum_j psi_j(s) exp(i theta xi t_j)|
  at given (s, xi) from a coefficient table, the quantity the expansion's
  sup_error bounds (gridder_gram.rs:143-163).
"""

import numpy as np

C_LIGHT = 299792458.0


def nm1(l, m):
    r2 = l * l + m * m
    return -r2 / (np.sqrt(np.maximum(1.0 - r2, 0.0)) + 1.0)


def gram_direct(uvw, freq, wgt, model, cell, flip_v=False, flip_w=False):
    """G m for an n x n model (float64); wgt (nrow, nchan) used as given."""
    n = model.shape[0]
    half = n // 2
    fv = -1.0 if flip_v else 1.0
    fw = -1.0 if flip_w else 1.0
    q = np.nonzero(model.ravel())[0]
    lp = (q // n - half) * cell
    mp = (q % n - half) * cell
    xp = nm1(lp, mp)
    vals = model.ravel()[q]
    i, j = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    l = ((i - half) * cell).ravel()
    mm = ((j - half) * cell).ravel()
    x = nm1(l, mm)
    out = np.zeros(n * n)
    foc = np.asarray(freq) / C_LIGHT
    for row in range(uvw.shape[0]):
        for ch in range(foc.size):
            u, v, w = uvw[row, 0] * foc[ch], fv * uvw[row, 1] * foc[ch], fw * uvw[row, 2] * foc[ch]
            wt = wgt[row, ch]
            if wt == 0.0:
                continue
            ph = 2.0 * np.pi * (u * (l[:, None] - lp[None, :]) + v * (mm[:, None] - mp[None, :])
                                + w * (x[:, None] - xp[None, :]))
            out += wt * (np.cos(ph) @ vals)
    return out.reshape(n, n)


def expansion_error(table, theta, s, xi):
    """|exp(i theta xi s) - sum_j psi_j(s) exp(i theta xi t_j)| with psi_j
    interpolated linearly in `table` ((bins + 1) x stencil)."""
    bins = table.shape[0] - 1
    stencil = table.shape[1]
    t = np.arange(stencil) - (stencil - 1) / 2.0
    f = (np.asarray(s) + 0.5) * bins
    b = np.clip(f.astype(int), 0, bins - 1)
    a = (f - b)[:, None]
    psi = table[b] + a * (table[b + 1] - table[b])
    xi = np.asarray(xi)
    exact = np.exp(1j * theta * xi * s)
    approx = (psi * np.exp(1j * theta * xi[:, None] * t[None, :])).sum(1)
    return np.abs(exact - approx)
