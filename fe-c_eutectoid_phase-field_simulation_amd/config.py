"""Input-file parser and temperature preparation (host side).

Restates read_config (pearl_serial.cpp:1600-1758) for the same strict
`key = value` files: '#' comments, duplicate and unknown keys are errors,
repeated FILL* lines form the ordered geometry recipe.  The temperature laws
use 80-bit long double exactly as the reference does:
  at_undercooling               :203-218  (linear_gp)
  at_phase_diagram_undercooling :236-276  (phase_diagram_fixed_curvature)
and the parsed values are checked bitwise against the reference parser in
tests/test_config.py.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

LD = np.longdouble
NAMES = ["alpha", "theta", "gamma"]


class ConfigError(RuntimeError):
    pass


def _validate_parabola(A, B, Cc, D):
    if not (math.isfinite(A) and math.isfinite(B) and math.isfinite(Cc) and math.isfinite(D)) or A <= 0 or D < 0:
        raise ConfigError("Parabola requires finite A>0, B, C, D>=0")


def at_undercooling(ref, slope, undercooling):
    """linear_gp law (:203-218).  ref=(A,B,C,D); slope=(q2,q1,q0)."""
    if not math.isfinite(undercooling) or undercooling < 0:
        raise ConfigError("undercooling must be finite and nonnegative")
    if not all(math.isfinite(s) for s in slope):
        raise ConfigError("nonfinite temperature derivative")
    _validate_parabola(*ref)
    if undercooling == 0:
        return tuple(ref)
    A, B, Cc = LD(ref[0]), LD(ref[1]), LD(ref[2])
    q2 = -1 / (4 * A)
    q1 = B / (2 * A)
    q0 = Cc - B * B / (4 * A)
    dT = -LD(undercooling)
    q2 = q2 + dT * LD(slope[0])
    q1 = q1 + dT * LD(slope[1])
    q0 = q0 + dT * LD(slope[2])
    if not (np.isfinite(q2) and np.isfinite(q1) and np.isfinite(q0)) or q2 >= 0:
        raise ConfigError("temperature law requires finite q2<0, q1, q0")
    A2 = -1 / (4 * q2)
    B2 = -q1 / (2 * q2)
    out = (float(np.float64(A2)), float(np.float64(B2)), float(np.float64(q0 + B2 * B2 / (4 * A2))), ref[3])
    _validate_parabola(*out)
    return out


def at_phase_diagram_undercooling(ref, ceq, slopes, undercooling):
    """phase_diagram_fixed_curvature law (:236-276).  ref: 3 x (A,B,C,D);
    ceq: alpha, theta, gamma contacts; slopes: alpha/ag, gamma/ag, theta/tg,
    gamma/tg."""
    if not math.isfinite(undercooling) or undercooling < 0:
        raise ConfigError("undercooling must be finite and nonnegative")
    for r in ref:
        _validate_parabola(*r)
    for c in ceq:
        if not math.isfinite(c) or c < 0 or c > 1:
            raise ConfigError("phase-diagram reference carbon must be in [0,1]")
    for s in slopes:
        if not math.isfinite(s):
            raise ConfigError("phase-diagram slope must be finite dc/dT")
    dT = -LD(undercooling)
    gA = LD(ref[2][0])
    gcE = LD(ceq[2])
    out = [list(r) for r in ref]
    for prod in (0, 1):
        pA = LD(ref[prod][0])
        pcE = LD(ceq[prod])
        pch = dT * LD(slopes[2 * prod])
        gch = dT * LD(slopes[2 * prod + 1])
        pc = pcE + pch
        gc = gcE + gch
        if not (np.isfinite(pc) and np.isfinite(gc)) or pc < 0 or pc > 1 or gc < 0 or gc > 1:
            raise ConfigError("phase-diagram extrapolated carbon outside [0,1]")
        if undercooling == 0:
            continue
        out[prod][1] = float(np.float64(LD(ref[prod][1]) + 2 * (gA * gch - pA * pch)))
        out[prod][2] = float(np.float64(LD(ref[prod][2]) + pA * pch * (2 * pcE + pch) - gA * gch * (2 * gcE + gch)))
        _validate_parabola(*out[prod])
    return [tuple(r) for r in out]


def _stod(s: str) -> float:
    """std::stod with the reference's full-consumption check."""
    try:
        v = float(s)
    except ValueError:
        raise ConfigError("invalid number")
    return v


@dataclass
class Config:
    nx: int = 0
    ny: int = 0
    dx: float = 0.0
    dy: float = 0.0
    thermo: list = field(default_factory=list)           # 3 x (A,B,C,D) at the run temperature
    reference_thermo: list = field(default_factory=list)
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
    raw: dict = field(default_factory=dict)

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


def _parse_fill(method: str, value: str, line: int):
    """parse_filling_operation (:1209-1245)."""
    def malformed():
        return ConfigError(f"Invalid {method} filling parameters at input line {line}: expected "
                           "{finite_number,...}, optionally followed by ';'")
    lst = value.strip(" \t\r\n")
    if lst.endswith(";"):
        lst = lst[:-1].strip(" \t\r\n")
    if len(lst) < 3 or lst[0] != "{" or lst[-1] != "}":
        raise malformed()
    items = lst[1:-1].split(",")
    out = []
    for it in items:
        it = it.strip(" \t\r\n")
        if not it:
            raise malformed()
        try:
            v = float(it)
        except ValueError:
            raise malformed()
        if not math.isfinite(v):
            raise malformed()
        out.append(v)
    return method, out


def read_config(path: str) -> Config:
    cfg = Config()
    raw = {}
    with open(path) as f:
        for ln, line in enumerate(f, 1):
            line = line.split("#", 1)[0].strip(" \t\r\n")
            if not line:
                continue
            if "=" not in line:
                raise ConfigError(f"missing = on line {ln}")
            key, val = line.split("=", 1)
            key, val = key.strip(" \t\r\n"), val.strip(" \t\r\n")
            if key.startswith("FILL"):
                cfg.filling.append(_parse_fill(key, val, ln))
                continue
            if not key or not val or key in raw:
                raise ConfigError("empty/duplicate key: " + key)
            raw[key] = val
    cfg.raw = raw
    used = set()

    def s(k, default=""):
        if k not in raw:
            return default
        used.add(k)
        return raw[k]

    def num(k, default=math.nan):
        v = s(k)
        if v == "":
            if not math.isfinite(default):
                raise ConfigError("missing required key: " + k)
            return default
        try:
            x = float(v)
        except ValueError:
            raise ConfigError("invalid number: " + k)
        if not math.isfinite(x):
            raise ConfigError("invalid number: " + k)
        return x

    def integer(k, default=-1):
        a = num(k, math.nan if default < 0 else float(default))
        if a != math.floor(a) or abs(a) > 1e8:
            raise ConfigError("invalid integer: " + k)
        return int(a)

    cfg.nx = integer("nx")
    cfg.ny = integer("ny")
    cfg.dx = num("dx")
    cfg.dy = num("dy", cfg.dx)
    cfg.eps, cfg.sigma, cfg.triple = num("epsilon"), num("sigma"), num("triple")
    cfg.Vm = num("Vm")
    cfg.tau = (num("tau_at"), num("tau_ag"), num("tau_tg"))
    cfg.thermo_mode = s("thermo_mode")
    if cfg.thermo_mode == "eutectoid_reference":
        cfg.T_eutectoid = num("T_eutectoid")
        cfg.undercooling = num("undercooling")
        cfg.max_undercooling = num("max_undercooling")
        cfg.temperature_law = s("temperature_law")
        if cfg.temperature_law not in ("linear_gp", "phase_diagram_fixed_curvature"):
            raise ConfigError("temperature_law must be linear_gp or phase_diagram_fixed_curvature")
        cfg.T = cfg.T_eutectoid - cfg.undercooling
        if (cfg.T_eutectoid <= 0 or cfg.T <= 0 or cfg.undercooling < 0 or cfg.max_undercooling < 0
                or cfg.undercooling > cfg.max_undercooling):
            raise ConfigError("temperature outside declared undercooling interval")
        cfg.reference_thermo = []
        cfg.thermo = [None, None, None]
        slopes_gp = []
        for i, n in enumerate(NAMES):
            ref = (num("Aeq_" + n), num("Beq_" + n), num("Ceq_" + n), num("D_" + n))
            cfg.reference_thermo.append(ref)
            if cfg.temperature_law == "linear_gp":
                sl = (num("dq2_dT_" + n), num("dq1_dT_" + n), num("dq0_dT_" + n))
                slopes_gp.append(sl)
                cfg.thermo[i] = at_undercooling(ref, sl, cfg.undercooling)
        cfg.temperature_slope = slopes_gp
        if cfg.temperature_law == "phase_diagram_fixed_curvature":
            cfg.ceq = (num("ceq_alpha"), num("ceq_theta"), num("ceq_gamma"))
            cfg.slopes = (num("slope_alpha_on_alpha_gamma"), num("slope_gamma_on_alpha_gamma"),
                          num("slope_theta_on_theta_gamma"), num("slope_gamma_on_theta_gamma"))
            cfg.thermo = at_phase_diagram_undercooling(cfg.reference_thermo, cfg.ceq, cfg.slopes, cfg.undercooling)
    elif cfg.thermo_mode == "direct_at_temperature":
        cfg.T = num("temperature")
        if cfg.T <= 0:
            raise ConfigError("temperature must be positive")
        cfg.thermo = []
        for n in NAMES:
            t = (num("A_" + n), num("B_" + n), num("C_" + n), num("D_" + n))
            _validate_parabola(*t)
            cfg.thermo.append(t)
    else:
        raise ConfigError("Set thermo_mode=eutectoid_reference or direct_at_temperature. "
                          "Legacy inputs require an explicit mode and temperature; no temperature is guessed.")
    cfg.dt = num("dt")
    cfg.dt_max = num("dt_max", cfg.dt)
    cfg.dt_min = num("dt_min", cfg.dt / 1048576.0)
    cfg.end_time = num("end_time")
    cfg.output_dt = num("output_dt")
    cfg.max_steps = integer("max_steps", 100000000)
    cfg.initialization = s("initialization", "lamella")
    cfg.mode = s("mode", "coupled")
    if cfg.mode not in ("coupled", "phase_only"):
        raise ConfigError("mode must be coupled or phase_only")
    cfg.seed_height = num("seed_height", cfg.ny * cfg.dy / 4)
    cfg.theta_columns = integer("theta_columns", cfg.nx // 2)
    cfg.radius = num("radius", min(cfg.nx * cfg.dx, cfg.ny * cfg.dy) / 4)
    cfg.phase_a = integer("phase_a", 0)
    cfg.phase_b = integer("phase_b", 2)
    if cfg.initialization not in ("lamella", "sharp_smooth", "flat", "circle", "microsim"):
        raise ConfigError("unknown initialization")
    if cfg.initialization == "microsim":
        if not cfg.filling:
            raise ConfigError("initialization=microsim requires at least one FILL operation")
    elif cfg.filling:
        raise ConfigError("FILL operations require initialization=microsim; no geometry is silently ignored")
    cfg.smooth_steps = integer("smooth_steps", 100)
    cfg.smooth_min_steps = integer("smooth_min_steps", cfg.smooth_steps)
    cfg.smooth_energy_rtol = num("smooth_energy_rtol", 1e-10)
    cfg.smooth_max_phase_change = num("smooth_max_phase_change", 0.02)
    if cfg.phase_a < 0 or cfg.phase_a > 2 or cfg.phase_b < 0 or cfg.phase_b > 2 or cfg.phase_a == cfg.phase_b:
        raise ConfigError("invalid phase pair")
    choices = sum(k in raw for k in ("initial_mu", "gamma_carbon", "mean_carbon"))
    if choices != 1:
        raise ConfigError("specify exactly one of initial_mu, gamma_carbon, mean_carbon")
    if "initial_mu" in raw:
        cfg.composition_mode = "mu"
        cfg.initial_mu = num("initial_mu")
    if "gamma_carbon" in raw:
        cfg.composition_mode = "gamma"
        cfg.initial_carbon = num("gamma_carbon")
        A, B = cfg.thermo[2][0], cfg.thermo[2][1]
        cfg.initial_mu = math.fma(2.0 * A, cfg.initial_carbon, B) if hasattr(math, "fma") else _fma(2.0 * A, cfg.initial_carbon, B)
    if "mean_carbon" in raw:
        cfg.composition_mode = "mean"
        cfg.initial_carbon = num("mean_carbon")
    cfg.max_phase_change = num("max_phase_change", .04)
    cfg.cg_rtol = num("cg_rtol", 1e-11)
    cfg.cg_atol = num("cg_atol", 1e-14)
    cfg.cg_max = integer("cg_max", 2000)
    cfg.closure_tol = num("closure_tol", 1e-10)
    cfg.energy_atol = num("energy_atol", 1e-13 * cfg.sigma * max(cfg.nx * cfg.dx, cfg.ny * cfg.dy))
    cfg.moving_window = integer("moving_window", 0) != 0
    cfg.shift_trigger = num("shift_trigger", .65 * cfg.ny * cfg.dy)
    cfg.shift_target = num("shift_target", .35 * cfg.ny * cfg.dy)
    A, B = cfg.thermo[2][0], cfg.thermo[2][1]
    default_cg = cfg.initial_carbon if cfg.composition_mode == "gamma" else (cfg.initial_mu - B) / (2.0 * A)
    cfg.shift_gamma_carbon = num("shift_gamma_carbon", default_cg)
    for k in raw:
        if k not in used:
            raise ConfigError("unknown key: " + k)
    if not (cfg.eps > 0) or not (cfg.sigma > 0) or cfg.triple < 0 or not all(t > 0 for t in cfg.tau) or not (cfg.Vm > 0):
        raise ConfigError("Invalid interface parameters")
    if (cfg.nx < 3 or cfg.ny < 2 or cfg.dx <= 0 or cfg.dy <= 0 or cfg.dt <= 0 or cfg.dt_min <= 0
            or cfg.dt_max < cfg.dt or cfg.dt_min > cfg.dt or cfg.end_time <= 0 or cfg.output_dt <= 0
            or cfg.max_steps < 1):
        raise ConfigError("invalid geometry/time bounds")
    if (cfg.theta_columns < 1 or cfg.theta_columns >= cfg.nx or cfg.seed_height <= 0
            or cfg.seed_height >= cfg.ny * cfg.dy):
        raise ConfigError("invalid seed geometry")
    if (cfg.max_phase_change <= 0 or cfg.max_phase_change > .2 or cfg.cg_rtol <= 0 or cfg.cg_atol <= 0
            or cfg.closure_tol <= 0 or cfg.energy_atol < 0):
        raise ConfigError("invalid numerical control")
    if (cfg.smooth_steps < 1 or cfg.smooth_min_steps < 0 or cfg.smooth_min_steps > cfg.smooth_steps
            or cfg.smooth_energy_rtol <= 0 or cfg.smooth_max_phase_change <= 0 or cfg.smooth_max_phase_change > .2):
        raise ConfigError("invalid smoothing control")
    if cfg.moving_window and (cfg.shift_target <= 0 or cfg.shift_trigger <= cfg.shift_target
                              or cfg.shift_trigger >= cfg.ny * cfg.dy or cfg.shift_gamma_carbon < 0
                              or cfg.shift_gamma_carbon > 1):
        raise ConfigError("invalid moving window settings")
    return cfg


def _fma(a: float, b: float, c: float) -> float:
    """Correctly rounded a*b+c (exact rational arithmetic, one rounding)."""
    from fractions import Fraction
    if not (math.isfinite(a) and math.isfinite(b) and math.isfinite(c)):
        return a * b + c
    return float(Fraction(a) * Fraction(b) + Fraction(c))
