"""The C oracle's restructured sweep (face states and face fluxes computed
once) against the literal per-cell restatement of the reference call graph
(oracle/literal.py) on tiny grids: bit for bit."""
from __future__ import annotations

import numpy as np
import pytest

from oracle import literal as Lit
from tests.fields import cons_field

FLUX = {"hllc": 0, "force": 1}
LIM = {"van_leer": 0, "superbee": 1}
FORM = {"cons": 0, "prim": 1}


@pytest.mark.parametrize("kind,flux,limiter,form,bc", [
    ("rough", "hllc", "van_leer", "cons", (0,)),
    ("rough", "force", "superbee", "prim", (1,)),
    ("fast", "hllc", "van_leer", "prim", (0,)),
    ("smooth", "force", "van_leer", "cons", (2,)),
    ("riemann", "hllc", "superbee", "cons", (0,)),
])
def test_literal_1d(oracle, kind, flux, limiter, form, bc):
    U = cons_field(kind, (14,), seed=21)
    s = oracle.scheme(limiter=LIM[limiter], form=FORM[form], flux=FLUX[flux], bc=bc)
    dtdh = 0.4 / oracle.max_wavespeed_grid(U, s)
    ref = oracle.sweep(U, s, dtdh, 0)
    lit = np.array(Lit.sweep(U.tolist(), 1, (14, 1), 0, dtdh, flux=flux, limiter=limiter,
                             form=form, bc=bc))
    assert np.array_equal(lit.view(np.uint64), ref.view(np.uint64))


@pytest.mark.parametrize("dim", [0, 1])
@pytest.mark.parametrize("kind,flux,limiter,form,bc", [
    ("rough", "hllc", "van_leer", "cons", (0, 0)),
    ("rough", "force", "van_leer", "prim", (1, 0)),
    ("fast", "hllc", "superbee", "cons", (0, 1)),
    ("smooth", "force", "superbee", "prim", (2, 2)),
])
def test_literal_2d(oracle, kind, flux, limiter, form, bc, dim):
    U = cons_field(kind, (7, 9), seed=5)
    s = oracle.scheme(limiter=LIM[limiter], form=FORM[form], flux=FLUX[flux], bc=bc)
    dtdh = 0.4 / oracle.max_wavespeed_grid(U, s)
    ref = oracle.sweep(U, s, dtdh, dim)
    lit = np.array(Lit.sweep(U.tolist(), 2, (9, 7), dim, dtdh, flux=flux, limiter=limiter,
                             form=form, bc=bc))
    assert np.array_equal(lit.view(np.uint64), ref.view(np.uint64))


def test_literal_planted_mistake_is_caught(oracle):
    """A dropped cross-dimension chi term in the literal restatement must
    change the answer on rough 2D data (the comparison has teeth)."""
    U = cons_field("rough", (7, 9), seed=5)
    s = oracle.scheme()
    dtdh = 0.4 / oracle.max_wavespeed_grid(U, s)
    ref = oracle.sweep(U, s, dtdh, 0)
    saved = Lit.evolve

    def evolve_1d_chi(it, mat, dtdh_, dim, face, limiter):
        class OneDim:
            def __init__(self, inner):
                self.inner = inner

            def __call__(self):
                st = self.inner()
                return st
        nd_saved = it().nd
        # chi over the sweep dimension only
        chi_v = 1.7976931348623157e308
        sign = Lit.signum(it.offset(1, dim)().pressure(mat) - it.offset(-1, dim)().pressure(mat))
        chi_v = Lit.std_min(chi_v, Lit.std_min(Lit.chi(it, mat, dim),
                                               Lit.chi(it.offset(sign, dim), mat, dim)))
        delta = Lit.scale(chi_v * 0.5, limiter.limit(it, mat, dim))
        u = it().v
        assert nd_saved == 2
        if face == +1:
            bev = Lit.State(Lit.add(u, delta))
            return Lit.State(Lit.add(bev.v, Lit.scale(0.5 * dtdh_, Lit.sub(
                Lit.State(Lit.sub(u, delta)).flux(mat, dim), bev.flux(mat, dim)))))
        bev = Lit.State(Lit.sub(u, delta))
        return Lit.State(Lit.add(bev.v, Lit.scale(0.5 * dtdh_, Lit.sub(
            bev.flux(mat, dim), Lit.State(Lit.add(u, delta)).flux(mat, dim)))))

    try:
        Lit.evolve = evolve_1d_chi
        lit = np.array(Lit.sweep(U.tolist(), 2, (9, 7), 0, dtdh))
    finally:
        Lit.evolve = saved
    assert not np.array_equal(lit, ref)
