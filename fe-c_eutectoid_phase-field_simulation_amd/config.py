"""Input files, read by the drop-in program's own reader.

read_config(path) returns the run settings that the host program's
key-schema reader (csrc/host/settings.hpp, exported by libpearl_input.so
through include/pearl_input.h) forms from an input file: the reference's
read_config (pearl_serial.cpp:1600-1758) with its defaults, checks and error
texts, the temperature laws in its long-double expressions (:186-276), and
the FILL admission of parse_filling_operation (:1209-1245).  This module only
binds that reader; it does not parse input files itself.  Values cross the
boundary as exact hexadecimal reals (tests/test_config.py checks them
bitwise against the reference parser).
"""
from __future__ import annotations

import ctypes as C
import json
import os
from dataclasses import dataclass, field

NAMES = ["alpha", "theta", "gamma"]
PKG_DIR = os.path.dirname(os.path.abspath(__file__))
INPUT_LIB = os.path.join(PKG_DIR, "libpearl_input.so")


class ConfigError(RuntimeError):
    pass


_lib = None


def lib():
    """libpearl_input.so (host code only; compiled here when it is absent)."""
    global _lib
    if _lib is None:
        if not os.path.exists(INPUT_LIB):
            import importlib.util
            spec = importlib.util.spec_from_file_location("pearl_amd_build", os.path.join(PKG_DIR, "build.py"))
            b = importlib.util.module_from_spec(spec)
            spec.loader.exec_module(b)
            b.build_input()
        L = C.CDLL(INPUT_LIB)
        dp = C.POINTER(C.c_double)
        L.pearl_read_input.restype = C.c_longlong
        L.pearl_read_input.argtypes = [C.c_char_p, C.c_char_p, C.c_longlong, C.c_char_p, C.c_int]
        L.pearl_input_fills.restype = C.c_int
        L.pearl_input_fills.argtypes = [C.c_char_p, C.c_void_p, C.c_int, C.POINTER(C.c_int), C.c_char_p, C.c_int]
        L.pearl_linear_gp_law.restype = C.c_int
        L.pearl_linear_gp_law.argtypes = [dp, dp, C.c_double, dp, C.c_char_p, C.c_int]
        L.pearl_phase_diagram_law.restype = C.c_int
        L.pearl_phase_diagram_law.argtypes = [dp, dp, dp, C.c_double, dp, C.c_char_p, C.c_int]
        _lib = L
    return _lib


def _doubles(values):
    v = [float(x) for x in values]
    return (C.c_double * len(v))(*v)


def _error(buf) -> ConfigError:
    return ConfigError(buf.value.decode(errors="replace"))


def at_undercooling(ref, slope, undercooling):
    """linear_gp law (:203-218).  ref=(A,B,C,D); slope=(dq2/dT, dq1/dT, dq0/dT)."""
    out, err = (C.c_double * 4)(), C.create_string_buffer(512)
    if lib().pearl_linear_gp_law(_doubles(ref), _doubles(slope), float(undercooling), out, err, len(err)) != 0:
        raise _error(err)
    return tuple(out)


def at_phase_diagram_undercooling(ref, ceq, slopes, undercooling):
    """phase_diagram_fixed_curvature law (:236-276).  ref: 3 x (A,B,C,D);
    ceq: alpha, theta, gamma contacts; slopes: alpha/ag, gamma/ag, theta/tg,
    gamma/tg."""
    out, err = (C.c_double * 12)(), C.create_string_buffer(512)
    flat = [v for r in ref for v in r]
    if lib().pearl_phase_diagram_law(_doubles(flat), _doubles(ceq), _doubles(slopes), float(undercooling), out, err,
                                     len(err)) != 0:
        raise _error(err)
    return [tuple(out[4 * p:4 * p + 4]) for p in range(3)]


@dataclass
class Config:
    path: str = ""
    nx: int = 0
    ny: int = 0
    dx: float = 0.0
    dy: float = 0.0
    thermo: list = field(default_factory=list)           # 3 x (A,B,C,D) at the run temperature
    reference_thermo: list = field(default_factory=list)  # 3 x (A,B,C,D) at the eutectoid (eutectoid_reference)
    temperature_slope: list = field(default_factory=list)  # linear_gp: 3 x (dq2/dT, dq1/dT, dq0/dT)
    ceq: tuple = ()                                      # phase_diagram_fixed_curvature contacts
    slopes: tuple = ()                                   # and slopes dc/dT
    eps: float = 0.0
    sigma: float = 0.0
    triple: float = 0.0
    tau: tuple = (0.0, 0.0, 0.0)
    Vm: float = 0.0
    thermo_mode: str = ""
    temperature_law: str = "linear_gp"
    T: float = 0.0
    T_eutectoid: float = 0.0
    undercooling: float = 0.0
    max_undercooling: float = 0.0
    dt: float = 0.0
    dt_max: float = 0.0
    dt_min: float = 0.0
    end_time: float = 0.0
    output_dt: float = 0.0
    max_steps: int = 100000000
    initialization: str = "lamella"
    mode: str = "coupled"
    seed_height: float = 0.0
    theta_columns: int = 0
    radius: float = 0.0
    phase_a: int = 0
    phase_b: int = 2
    smooth_steps: int = 100
    smooth_min_steps: int = 100
    smooth_energy_rtol: float = 1e-10
    smooth_max_phase_change: float = 0.02
    composition_mode: str = ""
    initial_mu: float = 0.0
    initial_carbon: float = 0.0
    max_phase_change: float = 0.04
    cg_rtol: float = 1e-11
    cg_atol: float = 1e-14
    cg_max: int = 2000
    closure_tol: float = 1e-10
    energy_rtol: float = 1e-11
    energy_atol: float = 0.0
    max_retries: int = 24
    check_energy: bool = True
    moving_window: bool = False
    shift_trigger: float = 0.0
    shift_target: float = 0.0
    shift_gamma_carbon: float = 0.0
    filling: list = field(default_factory=list)          # [(method, [params...])]
    physics_signature: str = ""

    def solver_params(self) -> dict:
        """Parameters in the PARAM_KEYS layout of the C ABI binding."""
        d = {}
        for i in range(3):
            d[f"A{i}"], d[f"B{i}"], d[f"C{i}"], d[f"D{i}"] = self.thermo[i]
        d.update(eps=self.eps, sigma=self.sigma, triple=self.triple, tau_at=self.tau[0], tau_ag=self.tau[1],
                 tau_tg=self.tau[2], Vm=self.Vm, dx=self.dx, dy=self.dy, nx=self.nx, ny=self.ny,
                 dt_max=self.dt_max, dt_min=self.dt_min, max_phase_change=self.max_phase_change,
                 cg_rtol=self.cg_rtol, cg_atol=self.cg_atol, closure_tol=self.closure_tol,
                 energy_rtol=self.energy_rtol, energy_atol=self.energy_atol, cg_max=self.cg_max,
                 max_retries=self.max_retries, check_energy=int(self.check_energy),
                 phase_only=int(self.mode == "phase_only"))
        return d


def read_config(path: str) -> Config:
    """The run settings of the input file at `path`; ConfigError with the
    reference's message when the file is not a valid input."""
    err = C.create_string_buffer(1024)
    size = 1 << 16
    while True:
        buf = C.create_string_buffer(size)
        n = lib().pearl_read_input(os.fsencode(path), buf, size, err, len(err))
        if n < 0:
            raise _error(err)
        if n < size:
            break
        size = n + 1
    d = json.loads(buf.value.decode())

    def real(v):
        return float.fromhex(v)

    def reals(v):
        return tuple(real(x) for x in v)

    cfg = Config(path=os.fspath(path))
    cfg.nx, cfg.ny = d["nx"], d["ny"]
    cfg.dx, cfg.dy = real(d["dx"]), real(d["dy"])
    cfg.eps, cfg.sigma, cfg.triple, cfg.Vm = (real(d[k]) for k in ("epsilon", "sigma", "triple", "Vm"))
    cfg.tau = reals(d["tau"])
    cfg.thermo_mode, cfg.temperature_law = d["thermo_mode"], d["temperature_law"]
    cfg.T, cfg.T_eutectoid = real(d["temperature"]), real(d["T_eutectoid"])
    cfg.undercooling, cfg.max_undercooling = real(d["undercooling"]), real(d["max_undercooling"])
    cfg.thermo = [reals(f) for f in d["thermo"]]
    if "at_eutectoid" in d:
        cfg.reference_thermo = [reals(f) for f in d["at_eutectoid"]]
    if "dg_dT" in d:
        cfg.temperature_slope = [reals(g) for g in d["dg_dT"]]
    if "c_eq" in d:
        cfg.ceq, cfg.slopes = reals(d["c_eq"]), reals(d["slopes"])
    for k in ("dt", "dt_max", "dt_min", "end_time", "output_dt", "seed_height", "radius", "smooth_energy_rtol",
              "smooth_max_phase_change", "initial_mu", "initial_carbon", "max_phase_change", "cg_rtol", "cg_atol",
              "closure_tol", "energy_atol", "shift_trigger", "shift_target", "shift_gamma_carbon"):
        setattr(cfg, k, real(d[k]))
    for k in ("max_steps", "theta_columns", "phase_a", "phase_b", "smooth_steps", "smooth_min_steps", "cg_max"):
        setattr(cfg, k, d[k])
    cfg.initialization, cfg.mode, cfg.composition_mode = d["initialization"], d["mode"], d["composition_mode"]
    cfg.moving_window = d["moving_window"] != 0
    cfg.filling = [(f["method"], [real(v) for v in f["p"]]) for f in d["fill"]]
    cfg.physics_signature = d["physics_signature"]
    return cfg


def sharp_label_ops(cfg: Config, fill_type):
    """The sharp labels of `cfg`'s input as an array of `fill_type`
    (pearl_fill) records and their count: the admitted FILL lines of
    initialization = microsim, or sharp_smooth's two boxes (the reader's
    sharp_label_ops, pearl_serial.cpp:1772-1786)."""
    err, count = C.create_string_buffer(1024), C.c_int(0)
    path = os.fsencode(cfg.path)
    if lib().pearl_input_fills(path, None, 0, C.byref(count), err, len(err)) != 0:
        raise _error(err)
    arr = (fill_type * max(1, count.value))()
    if lib().pearl_input_fills(path, arr, count.value, C.byref(count), err, len(err)) != 0:
        raise _error(err)
    return arr, count.value
