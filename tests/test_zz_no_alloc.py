"""s355_gradient_noalloc, the gradient entry point that allocates nothing
(its limiter scratch lives in a fixed per-GPU array), against s355_gradient
through the C ABI: bit for bit with the MLP and WENO limiters (the two that
use scratch) and with Barth, on two calls in a row, the second on a smaller
mesh, and a mesh above the scratch maximum is refused before any launch."""
import numpy as np
import pytest

import sysu_cfd_parallel_amd as s
from helpers import params, push, random_state
from sysu_cfd_parallel_amd._abi import FIELD_ID

pytestmark = pytest.mark.gpu


def limited_gradients(case, limiter, seed):
    p = params()
    p.limiter = limiter
    d = s.Solver(case, p)
    z = d.sizes
    st = random_state(z.n_cell, z.n_bnd, z.n_face, seed)
    push(d, {k: st[k] for k in ("U", "P", "BU", "BPRE")})
    out = {}
    for name, fn in (("alloc", d.lib.s355_gradient), ("noalloc", d.lib.s355_gradient_noalloc)):
        for f, b, g in (("P", "BPRE", "DPDX"), ("U", "BU", "DUDX")):
            d.set(g, np.full(3 * z.n_cell, np.nan))
            rc = fn(d.ctx, FIELD_ID[f], FIELD_ID[b], FIELD_ID[g])
            assert rc == 0, d.lib.s355_last_error()
            d.call("synchronize")
            out[(name, g)] = d.get(g)
    d.close()
    return out


@pytest.mark.parametrize("limiter", [2, 3, 1])
def test_noalloc_gradient_equals_gradient_twice(limiter):
    # first a larger jittered mixed hex/tet mesh, then a smaller one: the
    # second call runs on the scratch the first left behind
    for case, seed in ((s.channel(14, 9, 10, jitter=0.15, split=0.2, seed=3), 5), (s.cavity(5, jitter=0.1, seed=2), 6)):
        out = limited_gradients(case, limiter, seed)
        for g in ("DPDX", "DUDX"):
            a, b = out[("alloc", g)], out[("noalloc", g)]
            assert np.all(np.isfinite(a))
            assert np.array_equal(a, b), (limiter, g)


def test_noalloc_gradient_refuses_meshes_above_the_scratch():
    # WENO needs 3 doubles per cell: 1,064,960 cells exceed 3 * 2^20 doubles
    p = params()
    p.limiter = 3
    d = s.Solver(s.channel(130, 64, 128), p)
    assert 3 * d.sizes.n_cell > 3 * (1 << 20)
    rc = d.lib.s355_gradient_noalloc(d.ctx, FIELD_ID["P"], FIELD_ID["BPRE"], FIELD_ID["DPDX"])
    assert rc == -1 and b"scratch" in d.lib.s355_last_error()
    d.close()
