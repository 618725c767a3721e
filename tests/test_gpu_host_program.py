"""The drop-in host program pearl_gpu against the reference program
pearl_serial (built from source under oracle/_ref) on the same input file:
same files, same step/rejection/PCG counts, fields within the FP64 bar, and
checkpoints that each program restarts from."""
import os
import subprocess
import sys

import numpy as np
import pytest

from conftest import ROOT, PKG_DIR, load_pkg

import ref_tool

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not os.path.exists(os.path.join(ROOT, "oracle", "_ref", "pearl_serial")),
                                 reason="reference program not built")]
EXE = os.path.join(PKG_DIR, "pearl_gpu")
REF = os.path.join(ROOT, "oracle", "_ref", "pearl_serial")


def say(*a):
    print(*a, flush=True)


def make_input(tmp, name, **ov):
    load_pkg()
    import importlib
    inputs = importlib.import_module("pearl_amd.inputs")
    p = os.path.join(tmp, name + ".in")
    overrides = {"smooth_steps": 20, "smooth_min_steps": 20}
    overrides.update(ov)
    inputs.colony_input(p, 48, 40, 2, 10, 4.5e-10, output_dt=1.5e-10, overrides=overrides)
    return p


def run(exe, args, timeout=180):
    r = subprocess.run([exe, *args], capture_output=True, text=True, timeout=timeout)
    return r.returncode, r.stdout, r.stderr


def read_status(path):
    return dict(line.split("=", 1) for line in open(path).read().splitlines() if "=" in line)


def load_checkpoint(inp, path):
    # the reference's own reader validates it; it runs in a child process
    # because this process uses the GPU
    return ref_tool.read_checkpoint(inp, path, os.path.dirname(path))


# Report entries that are sums over the cells: the program sums them on
# the device in the blocked order, the reference row by row.
REPORT_SUMS = {"smooth_energy_initial": 1e-12, "smooth_energy_final": 1e-12}


def same_report(path_ref, path_gpu):
    """initialization_report.txt: every entry equal, the blocked sums within
    1e-12 relative of the reference's sequential ones, and the final relative
    energy change (a quotient of two such sums' difference) within 1e-13."""
    a, b = read_status(path_ref), read_status(path_gpu)
    assert sorted(a) == sorted(b)
    for k in a:
        if k in REPORT_SUMS:
            assert abs(float(a[k]) - float(b[k])) <= REPORT_SUMS[k] * abs(float(a[k])), (k, a[k], b[k])
        elif k == "smooth_final_relative_energy_change":
            assert abs(float(a[k]) - float(b[k])) <= 1e-13, (k, a[k], b[k])
        else:
            assert a[k] == b[k], (k, a[k], b[k])


def test_program_matches_reference(tmp_path):
    tmp = str(tmp_path)
    inp = make_input(tmp, "colony")
    say("host program: reference run")
    rc, out_r, err_r = run(REF, [inp, os.path.join(tmp, "ref")])
    assert rc == 0, err_r
    say("host program: gpu run")
    rc, out_g, err_g = run(EXE, [inp, os.path.join(tmp, "gpu")])
    assert rc == 0, err_g
    ref, gpu = os.path.join(tmp, "ref"), os.path.join(tmp, "gpu")
    assert sorted(os.listdir(ref)) == sorted(os.listdir(gpu))
    for f in ["Input.in", "thermodynamics.txt", "COMPLETED.txt"]:
        assert open(os.path.join(ref, f)).read() == open(os.path.join(gpu, f)).read(), f
    same_report(os.path.join(ref, "initialization_report.txt"), os.path.join(gpu, "initialization_report.txt"))
    for f in sorted(os.listdir(ref)):
        if f.endswith(".xmf"):
            assert open(os.path.join(ref, f)).read() == open(os.path.join(gpu, f)).read(), f
    # diagnostics: same rows; sums agree to the summation-order bar
    dr = np.genfromtxt(os.path.join(ref, "diagnostics.csv"), delimiter=",", names=True)
    dg = np.genfromtxt(os.path.join(gpu, "diagnostics.csv"), delimiter=",", names=True)
    assert dr.shape == dg.shape
    for k in ["time", "step", "dt_next", "cg_last", "cg_peak", "rejected", "shift_distance", "c_min", "c_max",
              "mu_min", "mu_max", "phi_sum_error"]:
        assert np.array_equal(dr[k], dg[k]), k
    assert np.allclose(dr["mean_carbon"], dg["mean_carbon"], rtol=1e-12, atol=0)
    assert np.allclose(dr["reduced_energy"], dg["reduced_energy"], rtol=1e-12, atol=0)
    # checkpoints: each loads through the reference reader (signature, closure, ledger)
    phi_r, c_r, mu_r, sc_r = load_checkpoint(inp, os.path.join(ref, "checkpoint.h5"))
    phi_g, c_g, mu_g, sc_g = load_checkpoint(inp, os.path.join(gpu, "checkpoint.h5"))
    assert list(sc_r[[0, 1, 6, 8]]) == list(sc_g[[0, 1, 6, 8]])
    assert np.abs(phi_r - phi_g).max() <= 1e-10
    assert np.abs(mu_r - mu_g).max() <= 1e-5
    # frame 0 (the initialized state) is bitwise identical
    f0r = load_checkpoint(inp, os.path.join(ref, "frame_000000.h5"))
    f0g = load_checkpoint(inp, os.path.join(gpu, "frame_000000.h5"))
    for a, b in zip(f0r[:3], f0g[:3]):
        assert np.array_equal(a, b)


def test_restart_both_ways(tmp_path):
    tmp = str(tmp_path)
    inp = make_input(tmp, "colony")
    inp_long = make_input(tmp, "colony_long")
    with open(inp_long) as f:
        text = f.read().replace("end_time = 4.5e-10", "end_time = 9e-10")
    with open(inp_long, "w") as f:
        f.write(text)
    say("restart: reference run")
    rc, _, err = run(REF, [inp, os.path.join(tmp, "ref")])
    assert rc == 0, err
    say("restart: gpu continues from the reference checkpoint")
    rc, _, err = run(EXE, [inp_long, os.path.join(tmp, "gpu_cont"), "--restart", os.path.join(tmp, "ref", "checkpoint.h5")])
    assert rc == 0, err
    say("restart: reference continues from its own checkpoint")
    rc, _, err = run(REF, [inp_long, os.path.join(tmp, "ref_cont"), "--restart", os.path.join(tmp, "ref", "checkpoint.h5")])
    assert rc == 0, err
    a = load_checkpoint(inp_long, os.path.join(tmp, "ref_cont", "checkpoint.h5"))
    b = load_checkpoint(inp_long, os.path.join(tmp, "gpu_cont", "checkpoint.h5"))
    assert list(a[3][[0, 6, 8]]) == list(b[3][[0, 6, 8]])
    assert np.abs(a[2] - b[2]).max() <= 1e-5
    say("restart: reference continues from the gpu checkpoint")
    rc, _, err = run(EXE, [inp, os.path.join(tmp, "gpu")])
    assert rc == 0, err
    rc, _, err = run(REF, [inp_long, os.path.join(tmp, "ref_from_gpu"), "--restart", os.path.join(tmp, "gpu", "checkpoint.h5")])
    assert rc == 0, err


def test_cli_errors(tmp_path):
    tmp = str(tmp_path)
    inp = make_input(tmp, "colony")
    os.makedirs(os.path.join(tmp, "exists"))
    for args in [[inp, os.path.join(tmp, "exists")], [inp, os.path.join(tmp, "x1"), "--px", "2"],
                 [inp, os.path.join(tmp, "x2"), "--bogus", "1"], [inp]]:
        rc_r, _, err_r = run(REF, args)
        rc_g, _, err_g = run(EXE, args)
        assert rc_r == rc_g == 1
        assert err_r.strip() == err_g.strip(), (args, err_r, err_g)


def test_max_steps_stop(tmp_path):
    tmp = str(tmp_path)
    inp = make_input(tmp, "colony")
    rc, _, err = run(EXE, [inp, os.path.join(tmp, "gpu"), "--max-steps", "3"])
    assert rc == 0, err
    st = read_status(os.path.join(tmp, "gpu", "STOPPED.txt"))
    assert st["status"] == "USER_STOPPED" and st["step"] == "3"


def test_gpu_init_flag_is_accepted(tmp_path):
    """Initialization always runs on the device; --gpu-init, which once
    selected that, is accepted and changes nothing."""
    tmp = str(tmp_path)
    inp = make_input(tmp, "colony")
    rc, _, err = run(EXE, [inp, os.path.join(tmp, "plain"), "--max-steps", "2"])
    assert rc == 0, err
    rc, _, err = run(EXE, [inp, os.path.join(tmp, "flag"), "--max-steps", "2", "--gpu-init"])
    assert rc == 0, err
    for f in ["frame_000000.h5", "checkpoint.h5"]:
        a = load_checkpoint(inp, os.path.join(tmp, "plain", f))
        b = load_checkpoint(inp, os.path.join(tmp, "flag", f))
        for x, y in zip(a[:3], b[:3]):
            assert np.array_equal(x, y)
    assert (open(os.path.join(tmp, "plain", "initialization_report.txt")).read() ==
            open(os.path.join(tmp, "flag", "initialization_report.txt")).read())


@pytest.mark.parametrize("init", ["lamella", "flat", "circle"])
def test_diffuse_initializations_match_reference(tmp_path, init):
    """The diffuse-profile initializations (the profiles evaluated on the
    host per column / row / cell with the C library's sine, the carbon
    set-up and the interface energy of the report on the device): frame 0
    bitwise the reference program's, the report within the summation-order
    bars, and the first steps with equal counts."""
    tmp = str(tmp_path)
    inp = make_input(tmp, init, initialization=init, theta_columns=9, seed_height=repr(14 * 2.79e-9),
                     radius=repr(11 * 2.79e-9), phase_a=1, phase_b=2)
    with open(inp) as f:
        lines = [ln for ln in f.read().splitlines() if not ln.startswith("FILL")]
    with open(inp, "w") as f:
        f.write("\n".join(lines) + "\n")
    say("diffuse initialization", init)
    rc, _, err = run(REF, [inp, os.path.join(tmp, "ref"), "--max-steps", "3"])
    assert rc == 0, err
    rc, _, err = run(EXE, [inp, os.path.join(tmp, "gpu"), "--max-steps", "3"])
    assert rc == 0, err
    a = load_checkpoint(inp, os.path.join(tmp, "ref", "frame_000000.h5"))
    b = load_checkpoint(inp, os.path.join(tmp, "gpu", "frame_000000.h5"))
    for x, y in zip(a[:3], b[:3]):
        assert np.array_equal(x, y)
    same_report(os.path.join(tmp, "ref", "initialization_report.txt"),
                os.path.join(tmp, "gpu", "initialization_report.txt"))
    sr = read_status(os.path.join(tmp, "ref", "STOPPED.txt"))
    sg = read_status(os.path.join(tmp, "gpu", "STOPPED.txt"))
    for k in ("step", "rejected", "CG_total"):
        assert sr[k] == sg[k], k


def test_gpu_init_fill_recipe_matches_reference(tmp_path):
    """--gpu-init on a FILL recipe with every method (tests/fill_recipes.py):
    the sharp labels the program writes (init_sharp.h5) are the reference
    program's bit for bit (pearl_fill_phases), and so are the smoothed
    phases and the initialized state (frame 0: pearl_smooth,
    pearl_init_composition), as every smoothing decision agrees on this
    input."""
    import importlib

    import fill_recipes
    load_pkg()
    inputs = importlib.import_module("pearl_amd.inputs")
    tmp = str(tmp_path)
    inp = inputs.write_input(os.path.join(tmp, "recipe.in"), 300, 170, 3e-10, 1.5e-10,
                             fill_recipes.recipe(300, 170, 0, "FILLCENTERBOX2D", overlays=8),
                             overrides={"smooth_steps": 40, "smooth_min_steps": 40})
    say("gpu-init recipe: reference run")
    rc, _, err = run(REF, [inp, os.path.join(tmp, "ref"), "--max-steps", "2"])
    assert rc == 0, err
    say("gpu-init recipe: gpu run")
    rc, _, err = run(EXE, [inp, os.path.join(tmp, "gpu"), "--max-steps", "2", "--gpu-init"])
    assert rc == 0, err
    for f in ["init_sharp.h5", "init_smoothed.h5"]:
        r = subprocess.run(["/opt/conda/bin/h5diff", os.path.join(tmp, "ref", f), os.path.join(tmp, "gpu", f)],
                           capture_output=True, text=True)
        assert r.returncode == 0, (f, r.stdout[-2000:])
    a = load_checkpoint(inp, os.path.join(tmp, "ref", "frame_000000.h5"))
    b = load_checkpoint(inp, os.path.join(tmp, "gpu", "frame_000000.h5"))
    for x, y in zip(a[:3], b[:3]):
        assert np.array_equal(x, y)
