"""Pins of the oracle's weighted least-squares gradient (grad_scheme 1, the
alternative to the reference's Green-Gauss Gradient, navier_gradient.cpp:12-48,
which has no least-squares form).  The least-squares fit over the face
neighbours is exact for a field linear in space on any mesh when the
boundary values are exact, where Green-Gauss with linear interpolation is
not (skewed and non-orthogonal cells); the fit is invariant to the field's
constant; and on a uniform hex mesh the symmetric stencil of an interior
cell makes it exact for quadratics."""
import numpy as np
import pytest

import sysu_cfd_parallel_amd as s
from helpers import params
from oracle.pyoracle import Oracle

MESHES = {
    "channel_split": lambda: s.channel(8, 6, 5, jitter=0.15, split=0.3, seed=4),
    "step_tets": lambda: s.step(2, upstream=2, downstream=4, span=2, jitter=0.1, turbulent=False),
    "cavity_jitter": lambda: s.cavity(6, jitter=0.15, seed=1),
}


def linear_gradient(o, g, c0=0.7):
    x = o.dbls("CELL_X").reshape(-1, 3)
    xf = o.dbls("FACE_X").reshape(-1, 3)[: o.sizes.n_bnd]
    o.set("P", c0 + x @ g)
    o.set("BPRE", c0 + xf @ g)
    o.gradient("P", "BPRE", "DPDX")
    return o.get("DPDX").reshape(-1, 3)


@pytest.mark.parametrize("name", sorted(MESHES))
def test_least_squares_exact_for_linear_fields(name):
    p = params()
    p.grad_scheme = 1
    o = Oracle(MESHES[name](), p)
    g = np.array([0.3, -1.7, 2.2])
    out = linear_gradient(o, g)
    assert np.max(np.abs(out - g)) < 1e-11 * np.abs(g).max()
    # the constant does not enter
    assert np.allclose(linear_gradient(o, g, c0=-40.0), out, rtol=0, atol=1e-11)


def test_green_gauss_is_not_exact_on_skewed_cells():
    # the contrast that makes the least-squares form worth having
    o = Oracle(MESHES["channel_split"](), params())
    g = np.array([0.3, -1.7, 2.2])
    assert np.max(np.abs(linear_gradient(o, g) - g)) > 1e-3


def test_least_squares_quadratic_field_second_order_on_uniform_mesh():
    # on a uniform hex mesh the symmetric stencil of an interior cell makes
    # the fit exact for quadratics too (central differences)
    p = params()
    p.grad_scheme = 1
    o = Oracle(s.cavity(6), p)
    x = o.dbls("CELL_X").reshape(-1, 3)
    xf = o.dbls("FACE_X").reshape(-1, 3)[: o.sizes.n_bnd]
    f = lambda y: y[:, 0] ** 2 + 2 * y[:, 1] * y[:, 2]  # noqa: E731
    o.set("P", f(x))
    o.set("BPRE", f(xf))
    o.gradient("P", "BPRE", "DPDX")
    out = o.get("DPDX").reshape(-1, 3)
    exact = np.stack([2 * x[:, 0], 2 * x[:, 2], 2 * x[:, 1]], axis=1)
    h = 1.0 / 6
    inner = np.all((x > h) & (x < 1 - h), axis=1)
    assert inner.sum() == 64
    assert np.max(np.abs(out[inner] - exact[inner])) < 1e-12
