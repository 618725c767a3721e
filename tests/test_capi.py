"""The C-ABI library loads and exports every entry point declared in
include/pearl_gpu.h (no GPU calls)."""
import os
import re
import ctypes

from conftest import ROOT, load_pkg


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "pearl_gpu.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(pearl_[a-z_0-9]+)\s*\(", text)))


def test_header_declares_expected_api():
    syms = declared_symbols()
    for s in ["pearl_create", "pearl_advance", "pearl_maybe_shift", "pearl_set_state", "pearl_get_state",
              "pearl_last_error", "pearl_destroy"]:
        assert s in syms


def test_library_exports_every_declared_symbol():
    pkg = load_pkg()
    lib = ctypes.CDLL(pkg.LIB_PATH)
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing
    assert sorted(pkg.EXPORTS) == declared_symbols()


def test_input_library_exports_every_declared_symbol():
    """libpearl_input.so (the input reader) exports what include/pearl_input.h declares."""
    load_pkg()
    import importlib
    cfg = importlib.import_module("pearl_amd.config")
    text = open(os.path.join(ROOT, "include", "pearl_input.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    syms = sorted(set(re.findall(r"\b(pearl_[a-z_0-9]+)\s*\(", text)))
    assert syms == ["pearl_input_fills", "pearl_linear_gp_law", "pearl_phase_diagram_law", "pearl_read_input"]
    lib = ctypes.CDLL(cfg.INPUT_LIB)
    assert not [s for s in syms if not hasattr(lib, s)]


def test_error_messages_match_reference_text():
    pkg = load_pkg()
    assert pkg.WHY == ["phase increment", "linear/closure", "composition/phase bounds", "mass conservation",
                       "free-energy increase"]


def test_binding_structs_match_the_header(tmp_path):
    """Every ctypes mirror of a C-ABI struct has the header's size and field
    offsets (a C probe compiled with gcc prints them)."""
    import shutil
    import subprocess
    pkg = load_pkg()
    gcc = shutil.which("gcc")
    assert gcc, "gcc is part of this image"
    structs = {"pearl_params": pkg.Params, "pearl_scalars": pkg.Scalars, "pearl_step_info": pkg.StepInfo,
               "pearl_diffusion_info": pkg.DiffusionInfo, "pearl_diag": pkg.Diag,
               "pearl_smooth_report": pkg.SmoothReport, "pearl_fill": pkg.Fill}
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "pearl_gpu.h"', "int main(void) {"]
    for cname, py in structs.items():
        lines.append(f'  printf("{cname} size %zu\\n", sizeof({cname}));')
        for fname, _ in py._fields_:
            lines.append(f'  printf("{cname} {fname} %zu\\n", offsetof({cname}, {fname}));')
    lines += ["  return 0;", "}"]
    src = tmp_path / "probe.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "probe"
    subprocess.run([gcc, "-std=c99", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split("\n")
    got = {tuple(l.split()[:2]): int(l.split()[2]) for l in out if l}
    for cname, py in structs.items():
        assert got[(cname, "size")] == ctypes.sizeof(py), cname
        for fname, _ in py._fields_:
            assert got[(cname, fname)] == getattr(py, fname).offset, (cname, fname)
    # method codes of pearl_fill in the binding's order
    codes = subprocess.run([gcc, "-std=c99", "-E", "-dM", "-I", os.path.join(ROOT, "include"),
                            os.path.join(ROOT, "include", "pearl_gpu.h")], capture_output=True, text=True,
                           check=True).stdout
    defined = {l.split()[1]: l.split()[2] for l in codes.splitlines() if l.startswith("#define PEARL_FILL_")}
    for i, name in enumerate(pkg.FILL_METHODS):
        assert defined["PEARL_FILL_" + name[4:]] == str(i), name
    assert defined["PEARL_FILL_PARAMS"] == str(len(pkg.Fill().p))


def test_argument_checks_before_any_device_call():
    """Entry points reject bad arguments before they touch a device, so these
    calls run without a GPU."""
    pkg = load_pkg()
    L = pkg.lib()
    g = ctypes.c_double(1.0)
    for nbytes, reps in ((15, 1), (0, 1), (-16, 1), (1 << 20, 0), ((1 << 35) + 16, 1)):
        assert L.pearl_stream_probe(0, nbytes, reps, ctypes.byref(g)) == pkg.ERR_ARGUMENT, (nbytes, reps)
    assert L.pearl_stream_probe(0, 1 << 20, 1, None) == pkg.ERR_ARGUMENT
    us = ctypes.c_double(1.0)
    for reads, writes, nx, ny, rows, reps in ((0, 1, 64, 8, 4, 1), (9, 1, 64, 8, 4, 1), (1, 0, 64, 8, 4, 1),
                                              (1, 9, 64, 8, 4, 1), (1, 1, 0, 8, 4, 1), (1, 1, 64, 8, 3, 1),
                                              (1, 1, 64, 8, 4, 0), (1, 1, 1 << 16, 1 << 16, 4, 1)):
        assert L.pearl_stream_mix(0, reads, writes, nx, ny, rows, reps, ctypes.byref(us),
                                  ctypes.byref(g)) == pkg.ERR_ARGUMENT
    assert L.pearl_stream_mix(0, 1, 1, 64, 8, 4, 1, None, ctypes.byref(g)) == pkg.ERR_ARGUMENT
    # an empty profile kind or a null context is rejected before any device call
    assert L.pearl_init_profile(None, 0, None, None, None, 0, 2) == pkg.ERR_ARGUMENT
    # a null context is an argument error for every context call
    assert L.pearl_fill_phases(None, None, 0) == pkg.ERR_ARGUMENT
    mu, carbon, valid = ctypes.c_double(0), ctypes.c_double(0), ctypes.c_int(0)
    assert L.pearl_init_composition(None, 0, 1.0, ctypes.byref(mu), ctypes.byref(carbon),
                                    ctypes.byref(valid)) == pkg.ERR_ARGUMENT


def test_domain_size_limit():
    """A domain whose three phase planes would overflow the kernels' 32-bit
    plane offsets is rejected before any device call."""
    pkg = load_pkg()
    import sys
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from conftest import load_golden
    pd = dict(load_golden("small_48x40")["pd"])
    h = ctypes.c_void_p()
    for nx, ny in ((40000, 20000), (715827, 1000)):  # 800 M and 716 M cells
        pd.update(nx=nx, ny=ny)
        P = pkg.params_from_dict(pd)
        assert pkg.lib().pearl_create(ctypes.byref(P), 3.1e4, pkg.FP64, 0, 0, ctypes.byref(h)) == pkg.ERR_ARGUMENT
        assert not h.value
