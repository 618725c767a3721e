"""Input-file generator for the lamellar-colony workloads.

Writes Pearl key=value input files (the format parsed by read_config,
pearl_serial.cpp:1600-1758) with the physical parameters of the reference's
default phase-diagram input (examples/fe_c_phase_diagram_10K_smoke.in) and a
FILLCUBE lamella recipe (pearl_serial.cpp:1399-1405) for the initial colony.
"""
from __future__ import annotations

import math
import random

# Physical parameters of the default 10 K phase-diagram input.
DEFAULT_PARAMETERS = [
    ("thermo_mode", "eutectoid_reference"),
    ("T_eutectoid", "999.68459774"),
    ("undercooling", "10"),
    ("temperature_law", "phase_diagram_fixed_curvature"),
    ("max_undercooling", "10"),
    ("Aeq_alpha", "4724862.6373624019"),
    ("Beq_alpha", "22964.296428571855"),
    ("Ceq_alpha", "-42255.03701292308"),
    ("D_alpha", "2e-9"),
    ("Aeq_theta", "100000000"),
    ("Beq_theta", "-49968654.66102279"),
    ("Ceq_theta", "6207741.2462954065"),
    ("D_theta", "2e-9"),
    ("Aeq_gamma", "172591.65731409981"),
    ("Beq_gamma", "19450.725112218221"),
    ("Ceq_gamma", "-42053.816371518253"),
    ("D_gamma", "1e-9"),
    ("ceq_alpha", "0.00088690859310575857"),
    ("ceq_theta", "0.25"),
    ("ceq_gamma", "0.034458832049305582"),
    ("slope_alpha_on_alpha_gamma", "-4.6282639461434345e-06"),
    ("slope_gamma_on_alpha_gamma", "-0.00031198492411304955"),
    ("slope_theta_on_theta_gamma", "0"),
    ("slope_gamma_on_theta_gamma", "0.00010753713136114516"),
    ("dx", "2.79e-9"),
    ("Vm", "1e-6"),
    ("epsilon", "14e-9"),
    ("sigma", "0.5"),
    ("triple", "10"),
    ("tau_at", "1.72402799784e8"),
    ("tau_ag", "1.72402799784e8"),
    ("tau_tg", "7.11828875478e9"),
    ("dt", "1.5e-11"),
    ("dt_max", "1.5e-11"),
    ("dt_min", "1e-16"),
    ("gamma_carbon", "0.034458832049305582"),
]


def lamella_fills(nx: int, pairs: int, strip_rows: int, theta_fraction: float = 1.0 / 8.0,
                  spacing_noise: float = 0.0, seed: int = 0):
    """FILLCUBE lines for `pairs` cementite+ferrite lamella pairs across the
    periodic width, in rows 0..strip_rows-1.  Each pair starts with its
    cementite plate.  With spacing_noise>0 the pair periods are drawn as
    f_i ~ U[1-noise, 1+noise] (seeded), normalised to fill nx exactly."""
    rng = random.Random(seed)
    if spacing_noise > 0:
        f = [rng.uniform(1 - spacing_noise, 1 + spacing_noise) for _ in range(pairs)]
    else:
        f = [1.0] * pairs
    tot = sum(f)
    widths = [int(round(nx * fi / tot)) for fi in f]
    widths[-1] += nx - sum(widths)
    fills = []
    x0 = 0
    for w in widths:
        th = max(1, int(round(w * theta_fraction)))
        fills.append(("FILLCUBE", f"{{1,{x0},0,0,{x0 + th - 1},{strip_rows - 1},0}}"))
        fills.append(("FILLCUBE", f"{{0,{x0 + th},0,0,{x0 + w - 1},{strip_rows - 1},0}}"))
        x0 += w
    return fills


def write_input(path: str, nx: int, ny: int, end_time: float, output_dt: float, fills,
                overrides: dict | None = None, extra: dict | None = None) -> str:
    """Write an input file with the default physics, nx*ny cells, the given
    FILL recipe (initialization=microsim) and optional key overrides."""
    keys = dict(DEFAULT_PARAMETERS)
    keys.update({"nx": str(nx), "ny": str(ny), "end_time": repr(float(end_time)),
                 "output_dt": repr(float(output_dt)), "initialization": "microsim"})
    for k, v in (overrides or {}).items():
        keys[k] = str(v)
    for k, v in (extra or {}).items():
        keys[k] = str(v)
    with open(path, "w") as f:
        f.write("# generated lamellar-colony input\n")
        for k, v in keys.items():
            f.write(f"{k} = {v}\n")
        for k, v in fills:
            f.write(f"{k} = {v}\n")
    return path


def colony_input(path: str, nx: int, ny: int, pairs: int, strip_rows: int, end_time: float,
                 output_dt: float | None = None, undercooling: float = 10.0, spacing_noise: float = 0.0,
                 seed: int = 0, overrides: dict | None = None) -> str:
    fills = lamella_fills(nx, pairs, strip_rows, spacing_noise=spacing_noise, seed=seed)
    ov = {"undercooling": repr(float(undercooling)),
          "max_undercooling": repr(float(max(10.0, math.ceil(undercooling))))}
    ov.update(overrides or {})
    return write_input(path, nx, ny, end_time, output_dt if output_dt else end_time * 10, fills, ov)


def ensemble(count: int = 64, seed: int = 7, widths=(256, 512, 1024, 2048), heights=(128, 256, 512, 1024),
             undercooling=(5.0, 40.0), pairs=(2, 16)):
    """Seeded ensemble of independent colonies: widths and heights drawn from
    the given sets, undercooling uniform in the given range, 2..16 lamella
    pairs, colony height a quarter of the domain.  The cementite width
    follows the lever rule at the run temperature (computed by the caller
    from the contacts of that temperature; see lever_rule_fraction)."""
    rng = random.Random(seed)
    out = []
    for i in range(count):
        out.append(dict(index=i, nx=rng.choice(widths), ny=rng.choice(heights),
                        undercooling=rng.uniform(*undercooling), pairs=rng.randint(*pairs)))
    return out


def lever_rule_fraction(c_mean: float, c_alpha: float, c_theta: float) -> float:
    """Cementite fraction of ferrite + cementite grown from austenite of mean
    carbon c_mean (lever rule between the product contacts)."""
    return (c_mean - c_alpha) / (c_theta - c_alpha)


def colony_workload(nx: int, ny: int, pairs: int, strip_rows: int, seed: int, undercooling: float, directory: str,
                    moving_window: int = 1, overrides: dict | None = None):
    """The lamellar-colony workload: writes its input file into `directory`
    and returns (path, config, phi[3,ny,nx], carbon, mu).  Pair spacings vary
    by +/-5 % (seeded), cementite:ferrite 1:7, colony rows [0, strip_rows);
    the phases are diffuse (initial.diffuse_colony), the carbon is the one of
    the common initial chemical potential and mu is uniform.  `overrides`
    adds input keys (e.g. shift_trigger, shift_target)."""
    import os

    import numpy as np

    from . import config, initial
    path = os.path.join(directory, "workload.in")
    fills = lamella_fills(nx, pairs, strip_rows, spacing_noise=0.05, seed=seed)
    keys = {"undercooling": repr(float(undercooling)),
            "max_undercooling": repr(float(max(10.0, math.ceil(undercooling)))),
            "moving_window": int(moving_window)}
    keys.update(overrides or {})
    write_input(path, nx, ny, 1e-6, 3e-10, fills, overrides=keys)
    cfg = config.read_config(path)
    phi = initial.diffuse_colony(nx, ny, cfg.dx, cfg.dy, cfg.eps, initial.plates_from_fills(fills), strip_rows)
    c = initial.carbon_at_mu(phi, cfg.thermo, cfg.initial_mu)
    mu = np.full((ny, nx), cfg.initial_mu)
    return path, cfg, phi, c, mu


def ensemble_workload(count: int, seed: int, directory: str):
    """The seeded batched-ensemble workload: `count` colonies from
    ensemble(count, seed) (widths 256..2048, heights 128..1024,
    undercooling U[5, 40] K, 2..16 lamella pairs), each with the cementite
    width of the lever rule between the product contacts of its run
    temperature, spacing +/-5 %, colony a quarter of the domain high.
    Returns one dict per domain: spec, cfg, phi[3,ny,nx], c, mu (the diffuse
    initial state at the input's initial mu)."""
    import os

    import numpy as np

    from . import config, initial
    doms = []
    for spec in ensemble(count, seed):
        nx, ny, uc, pairs = spec["nx"], spec["ny"], spec["undercooling"], spec["pairs"]
        # ferrite contact at the run temperature (the default input's linear law)
        c_alpha = 0.00088690859310575857 + (-4.6282639461434345e-06) * (-uc)
        f_theta = lever_rule_fraction(0.034458832049305582, c_alpha, 0.25)
        strip = ny // 4
        fills = lamella_fills(nx, pairs, strip, theta_fraction=f_theta, spacing_noise=0.05,
                              seed=seed * 1000 + spec["index"])
        path = os.path.join(directory, f"dom{spec['index']}.in")
        write_input(path, nx, ny, 1e-6, 1e-6, fills,
                    overrides={"undercooling": repr(float(uc)), "max_undercooling": "40"})
        cfg = config.read_config(path)
        phi = initial.diffuse_colony(nx, ny, cfg.dx, cfg.dy, cfg.eps, initial.plates_from_fills(fills), strip)
        c = initial.carbon_at_mu(phi, cfg.thermo, cfg.initial_mu)
        mu = np.full((ny, nx), cfg.initial_mu)
        doms.append(dict(spec=spec, cfg=cfg, phi=phi, c=c, mu=mu))
    return doms
