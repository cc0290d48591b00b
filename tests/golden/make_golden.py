"""Writes the golden fixtures under tests/golden/ as JSON data.

Every vector is transcribed from a reference artefact; nothing here runs the
reference.  Sources (paths relative to the reference root):

 state_kat.json    tests/state_tests_host.cpp:34-194 -- conservative/primitive
                   state accessors, prim<->cons round trips and the
                   conservative vs primitive Euler flux, checked with
                   EXPECT_EQ (or 1e-15 where the test says so).  2D states with
                   one additional component, IdealGas gamma = 1.4.
 boundary_kat.json tests/solver_tests_host.cpp:28-103 -- BoundarySetter mirror
                   mapping for transmissive and reflective ends of an 8-cell
                   1D array whose velocities are 0..7.
 flux_delta_kat.json
                   FaceFlux<MHReconstructor<VanLeer<cons>>, Flux,
                   IdealGas<double>>::flux_delta evaluated by the compiled
                   reference headers on the host at a Sod interface: cells
                   0..7 = (rho, p, u) (1, 1, 0), cells 8..15 = (0.125, 0.1, 0),
                   dtdh = 0.1, cell 7.  Values as printed by that evaluation
                   (17 significant digits).
 toro_cases.json   validation/1d_toro_case_{1..5}_validation.cu -- the five
                   Toro shock-tube set-ups (membrane, left/right states, end
                   time, resolution, CFL).
"""
from __future__ import annotations

import json
import os

HERE = os.path.dirname(os.path.abspath(__file__))


def state_kat():
    return {
        "gamma": 1.4,
        "ndim": 2,
        "nadd": 1,
        "cases": [
            {   # can_create_cons_state / can_convert_conservative_to_primitive
                "name": "cons_to_prim_round_trip",
                "cons": [1.0, 3.0, 2.0, 2.5, 3.0],
                "expect_prim_exact": {"rho": 1.0, "v_x": 2.0, "v_y": 2.5, "add": 3.0},
                "expect_round_trip_exact": [1.0, 3.0, 2.0, 2.5, 3.0],
            },
            {   # can_convert_primitive_to_conservative
                "name": "prim_to_cons_round_trip",
                "prim": [1.0, 2.0, 3.0, 4.0, 5.0],
                "expect_back_exact": {"rho": 1.0, "v_x": 3.0, "v_y": 4.0, "add": 5.0},
                "pressure_tolerance": 1e-15,
            },
            {   # primitive_and_conservative_fluxes_are_the_same
                "name": "flux_x",
                "rho": 1.0, "v_x": 2.0, "v_y": 2.5, "energy": 3.0, "add": 3.0,
                "dim": 0,
                "expect_exact": "rho*v_x, v_x*(E+p), rho*v_x*v_x+p, rho*v_x*v_y, rho*v_x*add"
                                " with p = (gamma-1)*(E - 0.5*rho*(v_x*v_x + v_y*v_y))",
            },
        ],
    }


def boundary_kat():
    vel = [float(i) for i in range(8)]
    return {
        "velocities": vel,
        "padding": 2,
        "transmissive": [3.0, 2.0, 2.0, 3.0, 4.0, 5.0, 5.0, 4.0],
        "reflective": [-3.0, -2.0, 2.0, 3.0, 4.0, 5.0, -5.0, -4.0],
    }


def flux_delta_kat():
    return {
        "left_prim": [1.0, 1.0, 0.0],
        "right_prim": [0.125, 0.1, 0.0],
        "cells_left": 8,
        "cells": 16,
        "dtdh": 0.1,
        "cell": 7,
        "gamma": 1.4,
        "limiter": "van_leer",
        "form": "conservative",
        "expect": {
            "force": [-2.21, -5.7019712000000009, 0.44855999999999996],
            "hllc": [-0.40261207906168767, -1.118424939519495, 0.47637703627371097],
        },
    }


def toro_cases():
    return [
        {"case": 1, "membrane": 0.3, "left": [1.0, 1.0, 0.75], "right": [0.125, 0.1, 0.0],
         "sim_time": 0.2, "resolution": 0.0025, "cfl": 0.9},
        {"case": 2, "membrane": 0.5, "left": [1.0, 0.4, -2.0], "right": [1.0, 0.4, 2.0],
         "sim_time": 0.15, "resolution": 0.01, "cfl": 0.9},
        {"case": 3, "membrane": 0.5, "left": [1.0, 1000.0, 0.0], "right": [1.0, 0.01, 0.0],
         "sim_time": 0.012, "resolution": 0.01, "cfl": 0.9},
        {"case": 4, "membrane": 0.4, "left": [5.99924, 460.894, 19.5975],
         "right": [5.99242, 46.0950, -6.19633], "sim_time": 0.035, "resolution": 0.01,
         "cfl": 0.9},
        {"case": 5, "membrane": 0.5, "left": [5.99924, 460.894, 19.5975],
         "right": [5.99242, 46.095, -6.19633], "sim_time": 0.035, "resolution": 0.005,
         "cfl": 0.9},
    ]


def main():
    files = {
        "state_kat.json": state_kat(),
        "boundary_kat.json": boundary_kat(),
        "flux_delta_kat.json": flux_delta_kat(),
        "toro_cases.json": toro_cases(),
    }
    for name, data in files.items():
        with open(os.path.join(HERE, name), "w") as f:
            json.dump(data, f, indent=1)
            f.write("\n")


if __name__ == "__main__":
    main()
