"""The settings front end (fluidity_amd.settings) on the CPU: the reference's
settings-file parse on its own validation settings file
(tests/golden/settings_2d_shock.txt, from validation/2d_shock_settings_new.txt)
and on hand-written strings whose split the reference's algorithm fixes
(setting/settings.hpp:49-90); the option choice lists and the mapping of the
runnable choices onto the C ABI's scheme enums."""
from __future__ import annotations

import os

import pytest

from fluidity_amd import native as N
from fluidity_amd.settings import (OPTIONS, OptionManager, Settings, UnsupportedOption,
                                   stod)

FIXTURE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                       "settings_2d_shock.txt")


def test_from_string_keeps_terminators():
    ss = Settings.from_string("a : 1, b : { c : 2, d : { e : 3 } }, f : 4")
    assert [s.name for s in ss] == ["a", "b", "f"]
    assert [s.complex for s in ss] == [False, True, False]
    assert ss[0].value == " 1,"
    assert ss[1].value == " { c : 2, d : { e : 3 } },"
    assert ss[2].value == " 4"  # unterminated: runs to the end of the text
    inner = Settings.from_string(ss[1].value)
    assert [s.name for s in inner] == ["c", "d"]
    assert inner[0].value == " 2,"
    assert inner[1].value == " { e : 3 } }"


def test_stod():
    assert stod(" 0.9,") == 0.9
    assert stod(" 1.6 }") == 1.6
    assert stod("\n 1e-3\n}") == 0.001
    assert stod("-.5x") == -0.5
    with pytest.raises(ValueError):
        stod(" ,")


def test_reference_settings_file():
    ss = Settings.from_file(FIXTURE)
    assert [s.name for s in ss] == ["simulator", "cfl", "shock", "ambient_state", "domain",
                                    "geometry"]
    assert [s.complex for s in ss] == [True, False, True, True, True, True]
    m = OptionManager()
    m.configure(ss)
    p = m.params
    assert p.cfl == 0.9 and p.resolution == 0.001
    assert p.dims == {"x": (0.0, 1.6), "y": (0.0, 1.0)}
    assert p.mach == 1.22 and p.shock_position == {"x": 0.1}
    assert p.ambient == {"rho": 1.0, "v_x": 0.0, "v_y": 0.0, "p": 1.0}
    # configure() leaves the options at their defaults, as the reference does
    assert m.values == {k: v[0] for k, v in OPTIONS.items()}
    m.configure_options(ss)
    assert m.values["flux_method"] == "force" and m.values["limiter"] == "van-leer"
    assert m.values["reconstructor"] == "muscl-hancock"
    s = m.scheme()
    assert (s.flux, s.limiter, s.form) == (N.FL_FLUX_FORCE, N.FL_LIMITER_VAN_LEER,
                                           N.FL_FORM_CONSERVATIVE)
    assert m.dimensions() == 1


def test_options_mapping_and_errors():
    m = OptionManager()
    with pytest.raises(UnsupportedOption, match="limiter"):
        m.scheme()  # default limiter "void", default reconstructor "none"
    m.set("limiter", "superbee").set("reconstructor", "muscl-hancock")
    m.set("flux_method", "hllc").set("limit_form", "primitive").set("dimensions", "two")
    s = m.scheme()
    assert (s.flux, s.limiter, s.form) == (N.FL_FLUX_HLLC, N.FL_LIMITER_SUPERBEE,
                                           N.FL_FORM_PRIMITIVE)
    assert m.dimensions() == 2
    with pytest.raises(ValueError):
        m.set("flux_method", "roe")
    assert not m.try_set_option("no_such_option", "x")
    m.set("flux_method", "richtmyer")
    assert m.scheme().flux == N.FL_FLUX_RICHTMYER
    m.set("flux_method", "lax-friedrichs")
    assert m.scheme().flux == N.FL_FLUX_LAX_FRIEDRICHS
    m.set("flux_method", "force").set("solve_method", "unsplit")
    with pytest.raises(UnsupportedOption, match="solve_method"):
        m.scheme()


def test_cfl_parameter_range():
    m = OptionManager()
    assert not m.try_set_param(Settings.from_string("cfl : 1.5,")[0])
    assert not m.try_set_param(Settings.from_string("cfl : { a : 1 },")[0])
    assert m.try_set_param(Settings.from_string("cfl : 0.5,")[0]) and m.params.cfl == 0.5
