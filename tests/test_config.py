"""The input reader (the drop-in program's, bound by config.py) and its
temperature laws against the reference parser (bitwise) and against the
published 10 K values."""
import glob
import os
import tempfile

import pytest

from conftest import ROOT, load_pkg

import importlib.util
import sys


def cfgmod():
    load_pkg()
    import importlib
    return importlib.import_module("pearl_amd.config")


def inputs_mod():
    load_pkg()
    import importlib
    return importlib.import_module("pearl_amd.inputs")


def test_phase_diagram_law_published_values():
    """NEW_FREE_ENERGY.md (reference) table at 10 K: A, B, C of alpha and theta
    (the published values are decimal roundings of the FP64 results)."""
    C = cfgmod()
    ref = [(4724862.6373624019, 22964.296428571855, -42255.03701292308, 2e-9),
           (100000000.0, -49968654.66102279, 6207741.2462954065, 2e-9),
           (172591.65731409981, 19450.725112218221, -42053.816371518253, 1e-9)]
    ceq = (0.00088690859310575857, 0.25, 0.034458832049305582)
    slopes = (-4.6282639461434345e-06, -0.00031198492411304955, 0.0, 0.00010753713136114516)
    out = C.at_phase_diagram_undercooling(ref, ceq, slopes, 10.0)
    assert out[0][0] == ref[0][0] and out[1][0] == ref[1][0]
    assert out[0][1] == pytest.approx(23603.858102865859, rel=2e-16, abs=0)
    assert out[0][2] == pytest.approx(-42293.428311015959, rel=2e-16, abs=0)
    assert out[1][1] == pytest.approx(-49969025.861257277, rel=2e-16, abs=0)
    assert out[1][2] == pytest.approx(6207753.8378329016, rel=2e-16, abs=0)
    assert out[2] == ref[2]
    # zero undercooling preserves every supplied bit
    assert C.at_phase_diagram_undercooling(ref, ceq, slopes, 0.0) == ref
    # initial mu from gamma_carbon (NEW_FREE_ENERGY.md: 31345.33897721396 J/mol),
    # mu = fma(2A, c, B) rounded once
    from fractions import Fraction
    mu = float(Fraction(2.0 * ref[2][0]) * Fraction(ceq[2]) + Fraction(ref[2][1]))
    assert mu == pytest.approx(31345.33897721396, rel=1e-15)
    # the law rejects what the reference rejects, with its text
    with pytest.raises(C.ConfigError, match="phase-diagram extrapolated carbon outside"):
        C.at_phase_diagram_undercooling(ref, ceq, slopes, 1e9)


def example_inputs():
    files = sorted(glob.glob("/root/reference/examples/*.in")) + ["/root/reference/Input.in"]
    return [f for f in files if os.path.exists(f)]


@pytest.mark.skipif(not os.path.exists("/root/reference"), reason="reference sources not present")
@pytest.mark.parametrize("path", example_inputs())
def test_parser_matches_reference_parser(path):
    import oracle as O
    if not O.ref_available():
        pytest.skip("oracle/_ref not built")
    C = cfgmod()
    try:
        R = O.Ref(path)
    except RuntimeError as e:
        with pytest.raises(C.ConfigError):
            C.read_config(path)
        return
    cfg = C.read_config(path)
    d = cfg.solver_params()
    for k in ["A0", "B0", "C0", "D0", "A1", "B1", "C1", "D1", "A2", "B2", "C2", "D2", "eps", "sigma", "triple",
              "tau_at", "tau_ag", "tau_tg", "Vm", "dx", "dy", "nx", "ny", "dt_max", "dt_min",
              "max_phase_change", "cg_rtol", "cg_atol", "closure_tol", "energy_rtol", "energy_atol",
              "cg_max", "max_retries", "check_energy", "phase_only"]:
        assert d[k] == R.pd[k], (k, d[k], R.pd[k])
    assert cfg.initial_mu == R.pd["initial_mu"]
    assert cfg.shift_trigger == R.pd["shift_trigger"]
    assert cfg.shift_target == R.pd["shift_target"]
    assert cfg.shift_gamma_carbon == R.pd["shift_gamma_carbon"]
    assert cfg.end_time == R.pd["end_time"] and cfg.output_dt == R.pd["output_dt"]
    assert int(cfg.moving_window) == R.pd["moving_window"]


@pytest.mark.skipif(not os.path.exists("/root/reference"), reason="reference sources not present")
@pytest.mark.parametrize("uc", [5.0, 15.0, 20.0, 40.0])
def test_undercooling_sweep_matches_reference(uc):
    import oracle as O
    if not O.ref_available():
        pytest.skip("oracle/_ref not built")
    C = cfgmod()
    I = inputs_mod()
    path = os.path.join(tempfile.mkdtemp(), "uc.in")
    I.colony_input(path, 64, 32, 2, 8, 1e-9, undercooling=uc)
    R = O.Ref(path)
    d = C.read_config(path).solver_params()
    for k in ["A0", "B0", "C0", "A1", "B1", "C1", "A2", "B2", "C2"]:
        assert d[k] == R.pd[k]


def test_parser_rejects_unknown_and_duplicate_keys(tmp_path):
    C = cfgmod()
    I = inputs_mod()
    p = str(tmp_path / "a.in")
    I.colony_input(p, 64, 32, 2, 8, 1e-9)
    with open(p, "a") as f:
        f.write("bogus_key = 1\n")
    with pytest.raises(C.ConfigError, match="unknown key: bogus_key"):
        C.read_config(p)
    I.colony_input(p, 64, 32, 2, 8, 1e-9)
    with open(p, "a") as f:
        f.write("nx = 64\n")
    with pytest.raises(C.ConfigError, match="empty/duplicate key: nx"):
        C.read_config(p)


def test_reader_reports_derived_settings(tmp_path):
    """initial_mu from gamma_carbon is the austenite parabola's fma(2A, c, B);
    defaults the reference computes (dy = dx, dt_min = dt / 2^20,
    shift_gamma_carbon) come back exactly; the physics signature is the
    one the drop-in program stores in its frames."""
    from fractions import Fraction
    C = cfgmod()
    I = inputs_mod()
    p = str(tmp_path / "g.in")
    I.colony_input(p, 64, 32, 2, 8, 1e-9)
    with open(p) as f:
        lines = [l for l in f if l.split("=")[0].strip() not in ("dy", "dt_min", "shift_gamma_carbon")]
    with open(p, "w") as f:
        f.writelines(lines)
    text = "".join(lines)
    cfg = C.read_config(p)
    assert cfg.dy == cfg.dx and cfg.dt_min == cfg.dt / 1048576.0
    A, B = cfg.thermo[2][0], cfg.thermo[2][1]
    want = cfg.initial_carbon if cfg.composition_mode == "gamma" else (cfg.initial_mu - B) / (2.0 * A)
    assert cfg.shift_gamma_carbon == want
    if cfg.composition_mode == "gamma":
        A, B = cfg.thermo[2][0], cfg.thermo[2][1]
        assert cfg.initial_mu == float(Fraction(2.0 * A) * Fraction(cfg.initial_carbon) + Fraction(B))
    assert cfg.physics_signature.startswith("pearlite-v0.4-reference;")
    assert repr(cfg.nx) in cfg.physics_signature
    assert "nx" in text


def test_reader_errors_are_the_reference_texts(tmp_path):
    C = cfgmod()
    I = inputs_mod()
    p = str(tmp_path / "e.in")
    I.colony_input(p, 64, 32, 2, 8, 1e-9)
    with open(p, "a") as f:
        f.write("FILLCUBE = {1, 2}\n")
    with pytest.raises(C.ConfigError, match="FILLCUBE: expected 7 parameters"):
        C.read_config(p)
    with pytest.raises(C.ConfigError, match="cannot open input"):
        C.read_config(str(tmp_path / "absent.in"))
