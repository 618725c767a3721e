"""The drop-in program's input handling against the reference program, on
the CPU (--thermo-only: read the input, write thermodynamics.txt, stop
before any device work).

1. Every example input of the reference: the same exit code, the same
   stderr and a byte-identical thermodynamics.txt.
2. Mutated inputs (2,634 of them: each key dropped, each key
   set to one of a dozen odd values, each key duplicated, an unknown key, a
   line without '=', and extra keys and FILL lines of every kind): the same
   exit code and the same error text, and byte-identical
   thermodynamics.txt where the input is accepted."""
import os
import subprocess

import pytest

from conftest import ROOT, PKG_DIR

REF = os.path.join(ROOT, "oracle", "_ref", "pearl_serial")
EXE = os.path.join(PKG_DIR, "pearl_gpu")
EXAMPLES = "/root/reference/examples"

pytestmark = pytest.mark.skipif(not (os.path.exists(REF) and os.path.exists(EXE) and os.path.isdir(EXAMPLES)),
                                reason="reference program, drop-in program or reference examples missing")

ODD_VALUES = ["abc", "1.5", "-3", "0", "1e400", "nan", "2e8", "7", "lamella", "phase_only", "{1,2,3}", "0x10"]
EXTRA_LINES = ["initial_mu = 31000", "mean_carbon = 0.03", "moving_window = 1", "shift_trigger = 1e-3",
               "FILLCUBE = {1,0,0,0,5,5,0}", "FILLCUBE = {1,0,0,0,5,5}", "FILLCUBE = {3,0,0,0,5,5,0}",
               "FILLCYLINDER = {0,10,10,0,0,4}", "FILLCENTERBOX2D = {0,10,0,10,1,1}", "FILLFOO = {1}",
               "FILLYJUNCTIONPEARLITE2D = {1,2}", "FILLCUBE = {1,0,0,0,5,5,0};", "FILLCUBE = {1,0,0,0,5,5,x}",
               "FILLPEARLITEELLIPSE2D = {10,10,30,0,4,5,0.3,0.125,1,1,0}",
               "FILLADHEREDPEARLITEARM2D = {10,10,30,40,4,0.125,3,4,1,1,0}",
               "FILLYJUNCTIONLAMELLAE2D = {0,64,0,40,30,20,90,210,330,3,3,5,1,0,1,2}",
               "initialization = microsim", "thermo_mode = direct_at_temperature", "temperature_law = foo",
               "undercooling = 100", "dt_min = 1", "smooth_min_steps = 1000", "phase_a = 2"]
BASES = ["fe_c_phase_diagram_10K_smoke.in", "fe_c_microsim_cube.in", "fe_c_smoke.in", "symmetric_eutectic.in"]


def thermo_only(exe, inp, out):
    r = subprocess.run([exe, inp, out, "--thermo-only"], capture_output=True, text=True, timeout=60)
    report = None
    if r.returncode == 0:
        with open(os.path.join(out, "thermodynamics.txt")) as f:
            report = f.read()
    return r.returncode, r.stderr.strip(), report


def variants():
    out = []
    for name in BASES:
        lines = open(os.path.join(EXAMPLES, name)).read().splitlines()
        keyed = [i for i, ln in enumerate(lines) if "=" in ln.split("#")[0]]
        for i in keyed:
            key = lines[i].split("#")[0].split("=", 1)[0].strip()
            out.append(lines[:i] + lines[i + 1:])
            out.extend(lines[:i] + [f"{key} = {v}"] + lines[i + 1:] for v in ODD_VALUES)
            out.append(lines + [lines[i]])
        out.append(lines + ["bogus_key = 3"])
        out.append(lines + ["just words"])
        out.extend(lines + [x] for x in EXTRA_LINES)
    return out


def test_examples_match_reference(tmp_path):
    names = sorted(f for f in os.listdir(EXAMPLES) if f.endswith(".in"))
    assert names
    for name in names:
        inp = os.path.join(EXAMPLES, name)
        a = thermo_only(REF, inp, str(tmp_path / ("r_" + name)))
        b = thermo_only(EXE, inp, str(tmp_path / ("g_" + name)))
        assert a == b, name


def test_mutated_inputs_match_reference(tmp_path):
    sample = variants()
    accepted = 0
    for k, lines in enumerate(sample):
        inp = str(tmp_path / f"v{k}.in")
        with open(inp, "w") as f:
            f.write("\n".join(lines) + "\n")
        a = thermo_only(REF, inp, str(tmp_path / f"r{k}"))
        b = thermo_only(EXE, inp, str(tmp_path / f"g{k}"))
        assert a == b, ("\n".join(lines), a[:2], b[:2])
        accepted += a[0] == 0
    assert 0 < accepted < len(sample)
