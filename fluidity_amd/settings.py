"""Settings files and runtime options of the reference, mapped to this
library's schemes.

``Settings.from_string`` / ``from_file`` parse the reference's settings text
(setting/settings.hpp:49-104): ``name : value`` pairs separated by commas,
values in braces nested.  The parse keeps the reference's behaviour: a
setting's value runs to and includes the comma or closing brace that ends
it, names lose spaces, braces and newlines, and numeric values are read as
``std::stod`` reads them (leading space, trailing text ignored).

``OptionManager`` holds the option choices of setting/*_option.hpp (the first
choice of each list is its default, as ``get_default_choice`` takes index 0)
and the parameters of setting/parameter/parameter_managers.hpp (ambient
state, cfl, domain, shock).  ``configure`` applies parameters only, as the
reference's does (setting/option_manager.hpp:170-194, where applying options
is commented out); ``configure_options`` also applies the ``simulator`` block
of a settings file.  ``create`` builds a ``simulator.Simulator`` for the
chosen options, raising ``UnsupportedOption`` for a choice the library does
not run.
"""
from __future__ import annotations

import re
from dataclasses import dataclass, field

from . import native as N


class UnsupportedOption(ValueError):
    """A valid option choice that this library does not run."""


@dataclass
class Setting:
    """setting/setting.hpp:28-76."""
    name: str = ""
    value: str = ""
    complex: bool = False

    def is_empty(self) -> bool:
        return not self.name or not self.value


class Settings:
    """setting/settings.hpp:30-104."""
    delim, sep, open_, close = ":", ",", "{", "}"

    @staticmethod
    def from_string(data: str) -> list[Setting]:
        settings: list[Setting] = []
        start = 0
        while True:
            match = data.find(Settings.delim, start)
            if match < 0:
                break
            s = Setting()
            s.name = re.sub(r"[ {\n]", "", data[start:match])
            match += 1
            start = match
            depth = 0
            while depth != -1:
                if match >= len(data):  # unterminated value: ends with the text
                    match += 1
                    break
                c = data[match]
                match += 1
                if c == Settings.open_:
                    depth += 1
                    s.complex = True
                if (c == Settings.sep and depth == 0) or c == Settings.close:
                    depth -= 1
            s.value = data[start:match]
            settings.append(s)
            start = match + 1
        return settings

    @staticmethod
    def from_file(path: str) -> list[Setting]:
        with open(path) as f:
            return Settings.from_string(f.read())


def stod(text: str) -> float:
    """std::stod: skips leading whitespace, reads the longest number prefix."""
    m = re.match(r"\s*([+-]?(?:\d+\.?\d*(?:[eE][+-]?\d+)?|\.\d+(?:[eE][+-]?\d+)?|inf|nan))",
                 text)
    if not m:
        raise ValueError(f"stod: no conversion for {text!r}")
    return float(m.group(1))


# option type -> choices (setting/*_option.hpp); the first is the default
OPTIONS: dict[str, tuple[str, ...]] = {
    "data_type": ("double", "float"),
    "dimensions": ("one", "two", "three"),
    "execution": ("gpu",),
    "flux_method": ("force", "hllc", "lax-friedrichs", "richtmyer"),
    "limit_form": ("conservative", "primitive"),
    "limiter": ("void", "linear", "van-leer", "superbee"),
    "material": ("ideal-gas",),
    "reconstructor": ("none", "muscl-hancock"),
    "solve_method": ("split", "unsplit"),
}

_RUNS = {
    "data_type": {"double", "float"},
    "dimensions": {"one", "two"},
    "execution": {"gpu"},
    "flux_method": {"force", "hllc", "lax-friedrichs", "richtmyer"},
    "limit_form": {"conservative", "primitive"},
    "limiter": {"van-leer", "superbee"},
    "material": {"ideal-gas"},
    "reconstructor": {"muscl-hancock"},
    "solve_method": {"split"},
}


@dataclass
class Parameters:
    """The parameters of setting/parameter/*.hpp that were set."""
    cfl: float | None = None
    resolution: float | None = None
    dims: dict[str, tuple[float, float]] = field(default_factory=dict)
    mach: float | None = None
    shock_position: dict[str, float] = field(default_factory=dict)
    ambient: dict[str, float] = field(default_factory=dict)


class OptionManager:
    """setting/option_manager.hpp:126-290 for this library's simulator."""

    def __init__(self):
        self.values = {k: v[0] for k, v in OPTIONS.items()}
        self.params = Parameters()

    # ---- options -------------------------------------------------------------
    def try_set_option(self, op_type: str, op_value: str) -> bool:
        """True when op_type names an option and op_value is one of its choices."""
        choices = OPTIONS.get(op_type)
        if choices is None or op_value not in choices:
            return False
        self.values[op_type] = op_value
        return True

    def set(self, op_type: str, op_value: str) -> "OptionManager":
        if not self.try_set_option(op_type, op_value):
            raise ValueError(f"{op_value!r} is not a choice of option {op_type!r}: "
                             f"{OPTIONS.get(op_type)}")
        return self

    # ---- parameters ------------------------------------------------------------
    def try_set_param(self, s: Setting) -> bool:
        """ParameterManager::try_set_param: each parameter takes a setting
        whose name contains the parameter's name."""
        if "cfl" in s.name:
            if s.complex:
                return False
            v = stod(s.value)
            if not 0.0 < v <= 1.0:
                return False
            self.params.cfl = v
            return True
        if "domain" in s.name:
            for prop in Settings.from_string(s.value):
                if prop.name == "resolution":
                    self.params.resolution = stod(prop.value)
                    continue
                dim = next((d for d in "xyz" if d in prop.name), None)
                if dim is None:
                    return False
                inner = Settings.from_string(prop.value)
                self.params.dims[dim] = (stod(inner[0].value), stod(inner[-1].value))
            return bool(self.params.resolution and self.params.resolution > 0 and
                        self.params.dims)
        if "shock" in s.name:
            for prop in Settings.from_string(s.value):
                if prop.name == "mach":
                    self.params.mach = stod(prop.value)
                    continue
                dim = next((d for d in "xyz" if d in prop.name), None)
                if dim is None:
                    continue
                self.params.shock_position[dim] = stod(prop.value)
            return True
        if "ambient_state" in s.name:
            for prop in Settings.from_string(s.value):
                if prop.name not in ("rho", "p", "v_x", "v_y", "v_z"):
                    return False
                self.params.ambient[prop.name] = stod(re.sub(r"[,} \n]", "", prop.value))
            return True
        return False

    def configure(self, settings: list[Setting]) -> None:
        """Applies the parameters among `settings` (the reference's configure)."""
        for s in settings:
            self.try_set_param(s)

    def configure_options(self, settings: list[Setting]) -> None:
        """Applies the parameters and the options of a ``simulator`` block."""
        self.configure(settings)
        for s in settings:
            if s.name == "simulator" and s.complex:
                for opt in Settings.from_string(s.value):
                    value = re.sub(r"[,} \n]", "", opt.value)
                    if not self.try_set_option(opt.name, value):
                        raise ValueError(f"invalid option {opt.name} : {value}")

    # ---- creation --------------------------------------------------------------
    def scheme(self):
        """The fl_scheme choices of the current options."""
        from .euler import Scheme
        for op, value in self.values.items():
            if value not in _RUNS[op]:
                raise UnsupportedOption(f"option {op} = {value} is not run by this library "
                                        f"(runs: {sorted(_RUNS[op])})")
        return Scheme(
            limiter=N.FL_LIMITER_SUPERBEE if self.values["limiter"] == "superbee"
            else N.FL_LIMITER_VAN_LEER,
            form=N.FL_FORM_PRIMITIVE if self.values["limit_form"] == "primitive"
            else N.FL_FORM_CONSERVATIVE,
            flux={"force": N.FL_FLUX_FORCE, "hllc": N.FL_FLUX_HLLC,
                  "lax-friedrichs": N.FL_FLUX_LAX_FRIEDRICHS,
                  "richtmyer": N.FL_FLUX_RICHTMYER}[self.values["flux_method"]])

    def dimensions(self) -> int:
        return {"one": 1, "two": 2, "three": 3}[self.values["dimensions"]]

    def create(self, device: int = 0):
        """A simulator.Simulator for the options, configured with the cfl and
        domain parameters that were set."""
        import torch
        from .simulator import Simulator
        scheme = self.scheme()
        dtype = torch.float32 if self.values["data_type"] == "float" else torch.float64
        sim = Simulator(dimensions=self.dimensions(), scheme=scheme, dtype=dtype, device=device)
        if self.params.cfl is not None:
            sim.configure_cfl(self.params.cfl)
        if self.params.resolution:
            sim.configure_resolution(self.params.resolution)
            for d, dim in enumerate("xyz"[:sim.dimensions]):
                if dim in self.params.dims:
                    sim.configure_dimension(d, *self.params.dims[dim])
        return sim
