"""The Gram operator test case of rang-core core/src/gridder_gram.rs:735-779:
xorshift64 from state 77, values (state >> 11) / 2^53 - 1/2; frequencies c
and 1.05 c; a cell of 8e-3 rad; uvw uniform in +-0.45/cell/1.05 (u, v) and
+-wmax (w); weights 0.2 + |value| normalised to sum to one."""

import numpy as np

C_LIGHT = 299792458.0


def xorshift_values(count, state=77):
    out = np.empty(count)
    s = state
    mask = (1 << 64) - 1
    for i in range(count):
        s ^= (s << 13) & mask
        s ^= s >> 7
        s ^= (s << 17) & mask
        out[i] = (s >> 11) / float(1 << 53) - 0.5
    return out, s


def gram_case(n, nrow, wmax, flip):
    vals, state = xorshift_values(3 * nrow)
    cell = 8e-3
    umax = 0.45 / cell
    idx = np.arange(3 * nrow)
    scale = np.where(idx % 3 == 2, wmax, umax / 1.05)
    uvw = (vals * 2.0 * scale).reshape(nrow, 3)
    raw, _ = xorshift_values(2 * nrow, state)
    raw = 0.2 + np.abs(raw)
    wgt = (raw / raw.sum()).reshape(nrow, 2)
    freq = np.array([C_LIGHT, 1.05 * C_LIGHT])
    return dict(uvw=uvw, freq=freq, wgt=wgt, n=n, cell=cell, flip=flip)


def model_of(n):
    """gridder_gram.rs:816-823: a corner (the largest x difference), the
    opposite corner, the centre and one more pixel."""
    m = np.zeros((n, n))
    m.flat[0] = 1.0
    m.flat[n * n - 1] = -0.5
    m.flat[(n // 2) * n + n // 2] = 0.75
    m.flat[5 * n + (n - 3)] = 0.25
    return m
