"""The w-expansion fit of the Gram operator (host code, no GPU) and the
oracle's direct-sum Gram against the oracle gridder.

Pins: the reference's own bounds (core/src/gridder_gram.rs:866-876:
stencil 8 at theta 1.21 sup error < 2e-4, stencil 12 at theta 1.0 < 3e-6,
Lebesgue constants < 20 and < 50); the expansion's error evaluated
independently at 20,000 random (s, xi) within 1.5 times the measured sup
(which is sampled on a grid); a planted mistake (the nodes' coefficients
swapped) that the same evaluation catches; and the direct-sum Gram equal to
vis2dirty(dirty2vis(.)) of the oracle gridder at epsilon 1e-12."""

import numpy as np
import pytest
from gram_cases import gram_case, model_of

from oracle import gram as OG
from oracle import oracle as O
from rang import gram


def test_expansion_bounds_of_the_reference():
    coarse = gram.WExpansion(8, 1.21, 1e-11)
    fine = gram.WExpansion(12, 1.0, 1e-13)
    assert coarse.sup_error < 2e-4 and fine.sup_error < 3e-6
    assert coarse.lebesgue < 20.0 and fine.lebesgue < 50.0


@pytest.mark.parametrize("stencil,theta,ridge", [(6, 1.2, 1e-13), (8, 1.21, 1e-11), (10, 1.0, 1e-13),
                                                 (12, 0.9, 1e-13)])
def test_expansion_error_at_random_points(stencil, theta, ridge):
    ex = gram.WExpansion(stencil, theta, ridge)
    rng = np.random.default_rng(stencil)
    s = rng.uniform(-0.5, 0.5, 20000)
    xi = rng.uniform(-1.0, 1.0, 20000)
    err = OG.expansion_error(ex.table, theta, s, xi)
    assert err.max() <= 1.5 * ex.sup_error
    # the fit is the point of the expansion: far better than one node
    assert ex.sup_error < 0.05 * theta
    # a planted mistake: the coefficients of the first two nodes swapped
    bad = ex.table.copy()
    bad[:, [0, 1]] = bad[:, [1, 0]]
    assert OG.expansion_error(bad, theta, s, xi).max() > 100 * ex.sup_error


def test_expansion_refuses_bad_arguments():
    with pytest.raises(ValueError, match="even"):
        gram.WExpansion(7, 1.0, 1e-13)
    with pytest.raises(ValueError, match="phase step"):
        gram.WExpansion(8, 3.5, 1e-13)


@pytest.mark.parametrize("flip", [False, True])
def test_direct_gram_is_the_gridders_normal_operator(flip):
    c = gram_case(16, 60, 300.0, flip)
    m = model_of(16)
    got = OG.gram_direct(c["uvw"], c["freq"], c["wgt"], m, c["cell"], flip_v=flip, flip_w=flip)
    kw = dict(pixsize_x=c["cell"], pixsize_y=c["cell"], epsilon=1e-12, flip_v=flip, flip_w=flip)
    vis, _ = O.dirty2vis(c["uvw"], c["freq"], m, None, **kw)
    img, _ = O.vis2dirty(c["uvw"], c["freq"], vis, c["wgt"], npix_x=16, npix_y=16, single=False, **kw)
    ref = img["dirty"]
    assert np.abs(got - ref).max() < 1e-10 * np.abs(ref).max()
    # symmetric: <a, G b> = <b, G a>
    b = np.zeros((16, 16))
    b.flat[40], b.flat[100] = 2.0, -1.0
    gb = OG.gram_direct(c["uvw"], c["freq"], c["wgt"], b, c["cell"], flip_v=flip, flip_w=flip)
    assert abs((m * gb).sum() - (b * got).sum()) < 1e-12
