"""The CPU oracle against the reference's own known-answer tests and the
values its compiled headers produced (tests/golden/*.json)."""
from __future__ import annotations

import json
import os

import numpy as np
import pytest

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    with open(os.path.join(GOLD, name)) as f:
        return json.load(f)


def test_state_kat_cons_to_prim(oracle):
    """tests/state_tests_host.cpp:107-139: EXPECT_EQ on the round trip."""
    kat = load("state_kat.json")
    s = oracle.scheme(gamma=kat["gamma"])
    case = kat["cases"][0]
    U = np.array(case["cons"])
    Wp = oracle.cons_to_prim(U, s, nd=2, na=1)
    e = case["expect_prim_exact"]
    assert Wp[0] == e["rho"] and Wp[2] == e["v_x"] and Wp[3] == e["v_y"] and Wp[4] == e["add"]
    back = oracle.prim_to_cons(Wp, s, nd=2, na=1)
    assert list(back) == case["expect_round_trip_exact"]
    # prim_state.pressure == cons_state.pressure exactly (:137)
    assert Wp[1] == oracle.pressure(U, s, nd=2, na=1)


def test_state_kat_prim_to_cons(oracle):
    """tests/state_tests_host.cpp:68-105."""
    kat = load("state_kat.json")
    s = oracle.scheme(gamma=kat["gamma"])
    case = kat["cases"][1]
    Wp = np.array(case["prim"])
    U = oracle.prim_to_cons(Wp, s, nd=2, na=1)
    back = oracle.cons_to_prim(U, s, nd=2, na=1)
    e = case["expect_back_exact"]
    assert back[0] == e["rho"] and back[2] == e["v_x"] and back[3] == e["v_y"]
    assert back[4] == e["add"]
    assert abs(back[1] - 2.0) < case["pressure_tolerance"]
    assert abs(oracle.pressure(U, s, nd=2, na=1) - Wp[1]) < case["pressure_tolerance"]


def test_state_kat_flux(oracle):
    """tests/state_tests_host.cpp:141-194: hand-computed conservative flux,
    and conservative flux == primitive flux, all EXPECT_EQ."""
    kat = load("state_kat.json")
    g = kat["gamma"]
    s = oracle.scheme(gamma=g)
    c = kat["cases"][2]
    rho, vx, vy, E, add = c["rho"], c["v_x"], c["v_y"], c["energy"], c["add"]
    p = (g - 1.0) * (E - 0.5 * rho * (vx * vx + vy * vy))
    U = np.array([rho, E, vx * rho, vy * rho, add])
    F = oracle.flux(U, s, 0, nd=2, na=1)
    assert F[0] == rho * vx
    assert F[1] == vx * (E + p)
    assert F[2] == rho * vx * vx + p
    assert F[3] == rho * vx * vy
    assert F[4] == rho * vx * add
    Wp = oracle.cons_to_prim(U, s, nd=2, na=1)
    Fp = oracle.flux(Wp, s, 0, nd=2, na=1, primitive=True)
    assert np.array_equal(F, Fp)


@pytest.mark.parametrize("kind", ["transmissive", "reflective"])
def test_boundary_kat(oracle, kind):
    """tests/solver_tests_host.cpp:28-103: ghost(-k-1) = setter(cell(k)),
    ghost(N+k) = setter(cell(N-1-k)); reflective negates the normal velocity."""
    kat = load("boundary_kat.json")
    bc = {"transmissive": oracle.BC_TRANSMISSIVE, "reflective": oracle.BC_REFLECTIVE}[kind]
    vel = kat["velocities"]
    P = kat["padding"]
    data = [np.array([1.0, 1.0, v]) for v in vel]  # rho = 1, so m = v exactly
    n = len(vel)
    # interior cells P..n-P-1; the setter fills ghosts from their mirrors
    for g in range(P):
        data[P - 1 - g] = oracle.apply_setter(data[P + g], 0, bc)
        data[n - P + g] = oracle.apply_setter(data[n - P - 1 - g], 0, bc)
    got = [d[2] / d[0] for d in data]
    assert got == kat[kind]
    assert all(d[0] == 1.0 for d in data)


@pytest.mark.parametrize("flux", ["force", "hllc"])
def test_flux_delta_kat(oracle, flux):
    """FaceFlux::flux_delta of the compiled reference at a Sod interface,
    reproduced bit for bit."""
    kat = load("flux_delta_kat.json")
    s = oracle.scheme(flux={"force": oracle.FLUX_FORCE, "hllc": oracle.FLUX_HLLC}[flux],
                      gamma=kat["gamma"])
    W = np.zeros((kat["cells"], 3))
    W[: kat["cells_left"]] = kat["left_prim"]
    W[kat["cells_left"]:] = kat["right_prim"]
    U = np.array([oracle.prim_to_cons(w, s) for w in W])
    got = oracle.flux_delta_at(U, s, kat["dtdh"], 0, kat["cell"])
    assert [float(x) for x in got] == kat["expect"][flux]


def test_lax_friedrichs_and_richtmyer_hand_values(oracle):
    """The standalone Lax-Friedrichs and Richtmyer fluxes
    (flux_method/flux_lax_friedrichs.hpp:62-78, flux_richtmyer.hpp:78-88) on a
    Sod-like pair at rest, gamma 1.4, dtdh 0.1, against values worked by hand:
      U_L = (1, 2.5, 0), U_R = (0.125, 0.25, 0) -> p_L = 1, p_R = 0.1,
      F_L = (0, 0, 1), F_R = (0, 0, 0.1);
      LF = 1/2 (F_L + F_R + 10 (U_L - U_R)) = (4.375, 11.25, 0.55);
      U* = 1/2 (U_L + U_R + 0.1 (F_L - F_R)) = (0.5625, 1.375, 0.045),
      u* = 0.08, p* = 0.4 (1.375 - 0.5 * 0.5625 * 0.0064) = 0.54928,
      RI = F(U*) = (0.045, 0.08 (1.375 + 0.54928), 0.045 * 0.08 + 0.54928).
    FORCE is their mean bit for bit, and each is consistent (F(U, U) = F(U))."""
    ul = np.array([1.0, 2.5, 0.0])
    ur = np.array([0.125, 0.25, 0.0])
    lf = oracle.numerical_flux(ul, ur, oracle.scheme(flux=oracle.FLUX_LF), 0.1, 0)
    ri = oracle.numerical_flux(ul, ur, oracle.scheme(flux=oracle.FLUX_RICHTMYER), 0.1, 0)
    fo = oracle.numerical_flux(ul, ur, oracle.scheme(flux=oracle.FLUX_FORCE), 0.1, 0)
    np.testing.assert_allclose(lf, [4.375, 11.25, 0.55], rtol=1e-15, atol=1e-15)
    np.testing.assert_allclose(ri, [0.045, 0.08 * (1.375 + 0.54928), 0.0036 + 0.54928],
                               rtol=1e-14, atol=1e-15)
    assert np.array_equal(fo, (lf + ri) * 0.5)
    u = np.array([0.7, 3.1, 0.4])
    F = oracle.flux(u, oracle.scheme(), 0)
    for fl in (oracle.FLUX_LF, oracle.FLUX_RICHTMYER):
        assert np.array_equal(oracle.numerical_flux(u, u, oracle.scheme(flux=fl), 0.3, 0), F)
