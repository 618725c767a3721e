"""initialize (pearl_serial.cpp:1761-1837) on the GPU: sharp FILL labels
(pearl_fill_phases) and the carbon set-up (pearl_init_composition), through
the C ABI, against the reference's own functions (oracle/_ref).

1. Labels of seeded recipes covering every FILL method, and sharp_smooth's
   two boxes: bitwise the reference's filling_phase_at (captured from its
   initialize at the init_sharp observer call), for several tile heights
   and widths that are not multiples of the 256-column tile; carbon and mu
   untouched.
2. Labels at the production sizes (4096x2048 with 64 lamella boxes,
   8192x4096 with 128): bitwise the numpy restatement of FILLCUBE
   (initial.sharp_labels, pinned to the reference by
   tests/test_initialize_labels.py through the same operations).
3. Carbon set-up from the reference's smoothed phases: with a given mu
   (initial_mu, gamma_carbon) c and mu are bitwise the reference's and the
   initial carbon agrees within the bound on two summation orders,
   2 (n-1) u sum|c| dx dy; with mean_carbon, mu
   agrees to 1e-12 relative (blocked against sequential sums of the
   intercept and susceptibility).
4. The whole device initialization (labels, pearl_smooth, carbon) against
   the reference's initialize: equal phases, carbon and mu where every
   smoothing decision agrees, which the equal step counts and energies show.
5. Argument and mixture errors.
6. The diffuse-profile initializations (pearl_init_profile): phases
   bitwise the reference's expressions of its lamella, flat and circle
   branches (:1767-1771, :1787-1791) from given profile values, at two tile
   heights and a width off the tile multiple; argument errors.  (The whole
   diffuse initialization of the drop-in program against the reference
   program's is tests/test_gpu_host_program.py.)"""
import importlib
import math
import os
import sys
import time

import numpy as np
import pytest

import fill_recipes as FR
import oracle
import ref_tool

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not oracle.ref_available(), reason="reference build (oracle/_ref) missing")]
DY = 2.79e-9


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def mod(name):
    return importlib.import_module("pearl_amd." + name)


def scalars(cfg):
    return [0.0, cfg.dt, 0.0, 0.0, 0.0, 0.0, 0, 0, 0]


def recipe_input(tmp_path, nx, ny, seed, whole_map, **ov):
    path = str(tmp_path / f"r_{nx}_{ny}_{seed}_{whole_map}.in")
    return mod("inputs").write_input(path, nx, ny, 1e-9, 1e-9, FR.recipe(nx, ny, seed, whole_map, overlays=8),
                                     overrides=ov)


@pytest.mark.parametrize("nx,ny,k", [(96, 64, 4), (300, 170, 1), (517, 203, 16)])
@pytest.mark.parametrize("whole_map", FR.WHOLE_MAP)
def test_fill_labels_equal_reference(pkg, tmp_path, nx, ny, k, whole_map):
    for seed in range(2):
        say("fill", nx, ny, whole_map, "seed", seed, "k", k)
        path = recipe_input(tmp_path, nx, ny, seed, whole_map)
        cfg = mod("config").read_config(path)
        want = ref_tool.sharp_labels(path, str(tmp_path))
        ops, n = pkg.fill_ops(cfg)
        c0 = np.linspace(0.01, 0.05, nx * ny)
        mu0 = np.full(nx * ny, cfg.initial_mu)
        with pkg.Solver(cfg.solver_params(), cfg.initial_mu, cells_per_thread=k) as s:
            s.set_state(np.full((3, nx * ny), 1.0 / 3.0), c0, mu0, scalars(cfg))
            s.fill_phases(ops, n)
            phi, c, mu, _ = s.get_state()
        assert np.array_equal(phi, want), int((phi != want).any(axis=0).sum())
        assert np.array_equal(c.reshape(-1), c0) and np.array_equal(mu.reshape(-1), mu0)


@pytest.mark.parametrize("nx,ny,theta_columns,rows", [(96, 64, 12, 10.5), (300, 170, 1, 33.0), (517, 203, 516, 0.7)])
def test_sharp_smooth_labels_equal_reference(pkg, tmp_path, nx, ny, theta_columns, rows):
    say("sharp_smooth", nx, ny, theta_columns, rows)
    path = mod("inputs").write_input(str(tmp_path / "s.in"), nx, ny, 1e-9, 1e-9, [],
                                     overrides={"initialization": "sharp_smooth", "theta_columns": theta_columns,
                                                "seed_height": repr(rows * DY)})
    cfg = mod("config").read_config(path)
    ops, n = pkg.fill_ops(cfg)
    with pkg.Solver(cfg.solver_params(), cfg.initial_mu) as s:
        s.fill_phases(ops, n)
        phi = s.get_state()[0]
    assert np.array_equal(phi, ref_tool.sharp_labels(path, str(tmp_path)))


@pytest.mark.parametrize("nx,ny,pairs,strip", [(4096, 2048, 32, 256), (8192, 4096, 64, 512)])
def test_fill_labels_at_production_size(pkg, tmp_path, nx, ny, pairs, strip):
    say("fill at", nx, ny)
    path = str(tmp_path / "colony.in")
    mod("inputs").colony_input(path, nx, ny, pairs, strip, 1e-6, spacing_noise=0.05, seed=nx)
    cfg = mod("config").read_config(path)
    ops, n = pkg.fill_ops(cfg)
    assert n == 2 * pairs
    with pkg.Solver(cfg.solver_params(), cfg.initial_mu) as s:
        t0 = time.perf_counter()
        s.fill_phases(ops, n)
        t1 = time.perf_counter()
        phi = s.get_state()[0]
    say(f"  pearl_fill_phases {1e3 * (t1 - t0):.1f} ms for {nx * ny} cells, {n} operations")
    want = mod("initial").sharp_labels(nx, ny, cfg.filling)
    assert np.array_equal(phi, want)


def sum_order_bound(c, cfg):
    """Bound on the difference of two summation orders of sum(c) dx dy:
    each recursive sum of n terms is within (n - 1) u sum|c| of the exact
    one (u = 2^-53)."""
    return 2.0 * (c.size - 1) * 2.0 ** -53 * float(np.sum(np.abs(c))) * cfg.dx * cfg.dy


def composition_input(tmp_path, mode):
    """A 96x64 recipe input whose composition is given by `mode`
    (initial_mu, gamma_carbon or mean_carbon; :1725-1740)."""
    path = recipe_input(tmp_path, 96, 64, 1, "FILLCENTERBOX2D", smooth_steps=30, smooth_min_steps=30)
    lines = [ln for ln in open(path).read().splitlines() if not ln.startswith("gamma_carbon")]
    value = {"initial_mu": "31340.25", "gamma_carbon": "0.034458832049305582", "mean_carbon": "0.05893"}[mode]
    lines.insert(1, f"{mode} = {value}")
    open(path, "w").write("\n".join(lines) + "\n")
    return path


@pytest.mark.parametrize("mode", ["initial_mu", "gamma_carbon", "mean_carbon"])
def test_composition_equals_reference(pkg, tmp_path, mode):
    say("composition", mode)
    path = composition_input(tmp_path, mode)
    cfg = mod("config").read_config(path)
    phi_r, c_r, mu_r, sc_r = ref_tool.initialize(path, str(tmp_path))
    n = cfg.nx * cfg.ny
    mean = cfg.composition_mode == "mean"
    with pkg.Solver(cfg.solver_params(), cfg.initial_mu) as s:
        s.set_state(phi_r.reshape(3, n), np.zeros(n), np.zeros(n), scalars(cfg))
        mu, carbon, valid = s.init_composition(pkg.COMPOSITION_MEAN if mean else pkg.COMPOSITION_MU,
                                               cfg.initial_carbon if mean else cfg.initial_mu)
        phi, c, m, _ = s.get_state()
    assert valid
    assert np.array_equal(phi, phi_r)
    say(f"  mu {mu!r} reference {sc_r[9]!r}; initial carbon {carbon!r} reference {sc_r[3]!r}")
    assert abs(carbon - sc_r[3]) <= sum_order_bound(c_r, cfg)
    if mean:
        assert abs(mu - sc_r[9]) <= 1e-12 * abs(sc_r[9])
        assert np.all(m == mu)
        assert np.max(np.abs(c - c_r)) <= 1e-12 * np.max(np.abs(c_r))
        # mean carbon of the result is the requested one
        assert abs(carbon / (cfg.nx * cfg.ny * cfg.dx * cfg.dy) - cfg.initial_carbon) <= 1e-12 * cfg.initial_carbon
    else:
        assert mu == sc_r[9] == cfg.initial_mu
        assert np.array_equal(m, mu_r) and np.array_equal(c, c_r)


@pytest.mark.parametrize("seed,whole_map", [(0, "FILLCENTERBOX2D"), (1, "FILLYJUNCTIONLAMELLAE2D")])
def test_device_initialize_equals_reference(pkg, tmp_path, seed, whole_map):
    say("device initialize", seed, whole_map)
    path = recipe_input(tmp_path, 300, 170, seed, whole_map, smooth_steps=40, smooth_min_steps=40)
    cfg = mod("config").read_config(path)
    phi_r, c_r, mu_r, sc_r = ref_tool.initialize(path, str(tmp_path))
    ops, n = pkg.fill_ops(cfg)
    with pkg.Solver(cfg.solver_params(), cfg.initial_mu) as s:
        s.fill_phases(ops, n)
        rep = s.smooth(cfg.smooth_steps, cfg.smooth_min_steps, cfg.smooth_energy_rtol, cfg.smooth_max_phase_change)
        mu, carbon, valid = s.init_composition(pkg.COMPOSITION_MU, cfg.initial_mu)
        phi, c, m, _ = s.get_state()
    say("  smoothing steps", rep["steps"], "energy", rep["energy_final"])
    assert valid and mu == sc_r[9]
    assert np.array_equal(phi, phi_r), float(np.max(np.abs(phi - phi_r)))
    assert np.array_equal(c, c_r) and np.array_equal(m, mu_r)
    assert abs(carbon - sc_r[3]) <= sum_order_bound(c_r, cfg)


def test_fill_argument_checks(pkg, tmp_path):
    path = recipe_input(tmp_path, 96, 64, 0, "FILLCENTERBOX2D")
    cfg = mod("config").read_config(path)
    n = cfg.nx * cfg.ny
    phi0 = mod("initial").sharp_labels(cfg.nx, cfg.ny, [("FILLCUBE", [1, 3, 2, 0, 40, 20, 0])]).reshape(3, n)
    with pkg.Solver(cfg.solver_params(), cfg.initial_mu) as s:
        s.set_state(phi0, np.full(n, 0.03), np.full(n, cfg.initial_mu), scalars(cfg))

        def rejected(edit, count=None):
            ops, k = pkg.fill_ops(cfg)
            edit(ops)
            with pytest.raises(pkg.PearlError) as e:
                s.fill_phases(ops, k if count is None else count)
            assert e.value.code == pkg.ERR_ARGUMENT
            return str(e.value)

        assert "invalid FILL operation 0" in rejected(lambda o: setattr(o[0], "method", 7))
        assert "invalid FILL operation 1" in rejected(lambda o: o[1].p.__setitem__(0, 3.0))
        assert "invalid FILL operation 2" in rejected(lambda o: o[2].p.__setitem__(5, math.nan))
        arm = next(i for i in range(len(cfg.filling)) if cfg.filling[i][0] == "FILLADHEREDPEARLITEARM2D")
        assert f"invalid FILL operation {arm}" in rejected(lambda o: o[arm].p.__setitem__(4, 0.0))
        assert "invalid FILL operation list" in rejected(lambda o: None, count=-1)
        # rejected calls leave the phases as they were
        assert np.array_equal(s.get_state()[0].reshape(3, n), phi0)
        with pytest.raises(pkg.PearlError) as e:
            s.init_composition(2, cfg.initial_mu)
        assert e.value.code == pkg.ERR_ARGUMENT
        with pytest.raises(pkg.PearlError) as e:
            s.init_composition(pkg.COMPOSITION_MU, math.inf)
        assert e.value.code == pkg.ERR_ARGUMENT


@pytest.mark.parametrize("mode", [0, 1])
def test_composition_reports_invalid_mixture(pkg, tmp_path, mode):
    """A non-finite phase fraction makes mixture() throw in the reference
    (:136); the set-up reports "Invalid mixture coefficients"."""
    path = recipe_input(tmp_path, 96, 64, 0, "FILLCENTERBOX2D")
    cfg = mod("config").read_config(path)
    n = cfg.nx * cfg.ny
    phi = np.zeros((3, n))
    phi[2] = 1.0
    phi[:, 1234] = math.nan
    with pkg.Solver(cfg.solver_params(), cfg.initial_mu) as s:
        s.set_state(phi, np.zeros(n), np.zeros(n), scalars(cfg))
        with pytest.raises(pkg.PearlError) as e:
            s.init_composition(mode, cfg.initial_mu if mode == 0 else 0.03)
    assert e.value.code == pkg.ERR_MIXTURE and "Invalid mixture coefficients" in str(e.value)


@pytest.mark.parametrize("k", [1, 4])
def test_init_profile_forms_the_reference_phases(pkg, tmp_path, k):
    nx, ny = 300, 170
    path = recipe_input(tmp_path, nx, ny, 0, "FILLCENTERBOX2D")
    cfg = mod("config").read_config(path)
    rng = np.random.default_rng(11)
    col, row, cell = rng.random(nx), rng.random(ny), rng.random((ny, nx))
    col[:7], row[:5] = 1.0, 0.0  # the profile's flat parts
    c0, mu0 = np.full((ny, nx), 0.03), np.full((ny, nx), cfg.initial_mu)
    with pkg.Solver(cfg.solver_params(), cfg.initial_mu, cells_per_thread=k) as s:
        s.set_state(np.zeros((3, ny, nx)), c0, mu0, scalars(cfg))
        s.init_profile(pkg.PROFILE_LAMELLA, col, row)
        phi, c, mu, _ = s.get_state()
        th, pr = col[None, :], row[:, None]
        want = np.stack([(1 - th) * pr, th * pr, np.broadcast_to(1 - pr, (ny, nx))])
        assert np.array_equal(phi, want)
        assert np.array_equal(c, c0) and np.array_equal(mu, mu0)
        s.init_profile(pkg.PROFILE_FLAT, row=row, phase_a=2, phase_b=0)
        phi = s.get_state()[0]
        assert np.array_equal(phi[2], np.broadcast_to(row[:, None], (ny, nx)))
        assert np.array_equal(phi[0], np.broadcast_to(1 - row[:, None], (ny, nx)))
        assert not phi[1].any()
        s.init_profile(pkg.PROFILE_CELLS, cell=cell, phase_a=1, phase_b=2)
        phi = s.get_state()[0]
        assert np.array_equal(phi[1], cell) and np.array_equal(phi[2], 1 - cell) and not phi[0].any()
        for kind, kw in ((3, dict(row=row)), (pkg.PROFILE_LAMELLA, dict(col=col)),
                         (pkg.PROFILE_FLAT, dict(row=row, phase_a=1, phase_b=1)),
                         (pkg.PROFILE_CELLS, dict(cell=cell, phase_a=0, phase_b=3)),
                         (pkg.PROFILE_FLAT, dict(row=np.where(np.arange(ny) == 9, np.nan, row)))):
            with pytest.raises(pkg.PearlError) as e:
                s.init_profile(kind, **kw)
            assert e.value.code == pkg.ERR_ARGUMENT and "invalid profile" in str(e.value)
        # the phases of the last valid call are still in place
        assert np.array_equal(s.get_state()[0][1], cell)
