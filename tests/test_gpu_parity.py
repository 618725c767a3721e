"""Parity of the HIP path (through the C ABI) with the oracle and with the
reference's own outputs (golden fixtures).

Bars:
 - per-cell work (propose_phi, max reductions, acceptance flags): bitwise
   equal to the reference;
 - everything: bitwise equal to the oracle run in the HIP path's blocked
   reduction order (same cells-per-thread layout);
 - against the reference's sequential sums: within the FP64 tolerance of the
   target (|dmu| <= 1e-5 J/mol, |dphi| <= 1e-10), with equal step counts,
   rejections and CG iterations on these fixtures.
"""
import sys

import numpy as np
import pytest

from conftest import load_golden

pytestmark = pytest.mark.gpu

CASES = ["small_48x40", "phase_only_48x40", "window_alpha_32x40", "colony_256x128", "tiny_3x2", "ragged_257x9"]
MU_TOL = 1e-5
PHI_TOL = 1e-10


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def make(pkg, g, k=0):
    s = pkg.Solver(g["pd"], float(g["sc0"][9]), cells_per_thread=k)
    s.set_state(g["phi0"], g["c0"], g["mu0"], g["sc0"][:9])
    return s


def blocked_params(O, g, k):
    return O.params_from_dict(g["pd"], red_mode=O.RED_BLOCKED, red_k=k)


@pytest.mark.parametrize("name", CASES)
@pytest.mark.parametrize("k", [1, 4])
def test_propose_phi_bitwise_vs_reference(pkg, name, k):
    g = load_golden(name)
    say("propose_phi", name, "k", k)
    with make(pkg, g, k) as s:
        ch, phit = s.propose_phi(float(g["dt_probe"]))
    assert ch == g["propose_change"]
    assert np.array_equal(phit, g["propose_phi"])


@pytest.mark.parametrize("k", [1, 2, 4, 16])
def test_propose_phi_out_of_range_numerators(pkg, O, k):
    """A patch of cells whose phase fractions are tiny (down to subnormal)
    makes face and driving-force numerators leave the fast constant
    quotient's range: those cells are computed again with IEEE quotients
    after the march, and their north face terms are not carried.  Bitwise
    the oracle's sequential propose_phi, which divides with IEEE quotients
    throughout."""
    g = load_golden("small_48x40")
    ny, nx = g["mu0"].shape
    phi = np.array(g["phi0"], dtype=np.float64).reshape(3, ny, nx)
    rng = np.random.default_rng(17)
    ys, xs = slice(5, 21), slice(3, 31)
    vals = np.array([0.0, 5e-324, 3e-310, 1e-200, 1e-160, 2e-155, 1e-100])
    t = rng.choice(vals, size=(2, 16, 28))
    phi[0, ys, xs], phi[1, ys, xs] = t[0], t[1]
    phi[2, ys, xs] = 1.0 - t[0] - t[1]
    # and single tiny cells next to ordinary interface cells
    for y, x in ((30, 10), (31, 10), (30, 11), (2, 40)):
        phi[:, y, x] = (1e-170, 3e-165, 1.0)
    say("propose_phi out-of-range numerators, k", k)
    with pkg.Solver(g["pd"], float(g["sc0"][9]), cells_per_thread=k) as s:
        s.set_state(phi, g["c0"], g["mu0"], g["sc0"][:9])
        ch, got = s.propose_phi(float(g["dt_probe"]))
    P = O.params_from_dict(g["pd"])
    want_ch, want, st = O.propose_phi(P, phi, np.asarray(g["mu0"], dtype=np.float64), float(g["dt_probe"]))
    assert ch == want_ch
    assert np.array_equal(got.reshape(want.shape), want)
    assert np.array_equal(np.signbit(got.reshape(want.shape)), np.signbit(want))


@pytest.mark.parametrize("k", [1, 4])
def test_propose_phi_change_reduction_edges(pkg, O, k):
    """propose_phi's largest change is folded by per-block atomic maxima on
    its bits, and a block whose maximum is +0 takes no atomic.  Edges of that
    reduction, each bitwise the oracle's sequential propose_phi:
     - every cell bulk austenite: no block changes, so no atomic is taken,
       and the change is exactly 0 with the phases unchanged;
     - one changed cell in the domain's last tile: a single block carries
       the maximum;
     - a chemical potential whose psi overflows in one cell: that cell's
       trial value is not finite, the change is +inf and the cell keeps its
       phases."""
    g = load_golden("small_48x40")
    ny, nx = g["mu0"].shape
    mu = np.asarray(g["mu0"], dtype=np.float64)
    dt = float(g["dt_probe"])
    P = O.params_from_dict(g["pd"])
    bulk = np.zeros((3, ny, nx))
    bulk[2] = 1.0
    one = bulk.copy()
    one[:, ny - 1, nx - 2] = (0.25, 0.125, 0.625)
    huge = mu.copy()
    huge[ny // 2, 5] = 1e300
    cases = [("bulk", bulk, mu, 0.0), ("one cell in the last tile", one, mu, None),
             ("psi overflow in one cell", np.asarray(g["phi0"], dtype=np.float64), huge, np.inf)]
    for label, phi, m, expect in cases:
        say("propose_phi reduction edge:", label, "k", k)
        want_ch, want, st = O.propose_phi(P, phi, m, dt)
        assert st == 0
        if expect is not None:
            assert want_ch == expect
        with pkg.Solver(g["pd"], float(g["sc0"][9]), cells_per_thread=k) as s:
            s.set_state(phi, g["c0"], m, g["sc0"][:9])
            ch, got = s.propose_phi(dt)
        assert ch == want_ch
        assert np.array_equal(got.reshape(want.shape), want)
        if label == "bulk":
            assert np.array_equal(got.reshape(phi.shape), phi)


def test_projection_failure(pkg, O):
    """A step so large that a trial value reaches 2^53: the projection finds
    no active set (pearl_serial.cpp:505 throws).  The oracle reports it
    (status 5, ORC_ERR_PROJECTION), the single propose entry returns
    PEARL_ERR_PROJECTION, and a whole step started by advance ends with the
    same error and leaves the state as it was."""
    g = load_golden("small_48x40")
    dt = float(g["dt_probe"]) * 1e20
    P = O.params_from_dict(g["pd"])
    _, _, st = O.propose_phi(P, np.asarray(g["phi0"], dtype=np.float64), np.asarray(g["mu0"], dtype=np.float64), dt)
    assert st == 5
    say("projection failure: single entry")
    with make(pkg, g) as s:
        with pytest.raises(pkg.PearlError) as e:
            s.propose_phi(dt)
        assert e.value.code == pkg.ERR_PROJECTION
    say("projection failure: advance")
    pd = dict(g["pd"])
    pd["dt_max"] = 2 * dt
    with pkg.Solver(pd, float(g["sc0"][9])) as s:
        s.set_state(g["phi0"], g["c0"], g["mu0"], g["sc0"][:9])
        sc = s.scalars()
        sc.dt = dt
        s.set_scalars(sc)
        before = s.get_state()
        with pytest.raises(pkg.PearlError) as e:
            s.advance(10 * dt)
        assert e.value.code == pkg.ERR_PROJECTION
        after = s.get_state()
        for x, y in zip(before[:3], after[:3]):
            assert np.array_equal(x, y)
        assert after[3].step == before[3].step and after[3].time == before[3].time


@pytest.mark.parametrize("name", CASES)
@pytest.mark.parametrize("k", [1, 2, 4, 8, 16])
def test_implicit_diffusion(pkg, O, name, k):
    g = load_golden(name)
    if g["pd"]["phase_only"]:
        pytest.skip("the reference calls implicit_diffusion only in coupled mode (pearl_serial.cpp:1134-1144)")
    say("implicit_diffusion", name, "k", k)
    with make(pkg, g, k) as s:
        d = s.implicit_diffusion(g["propose_phi"], float(g["dt_probe"]))
    P = blocked_params(O, g, k)
    o = O.implicit_diffusion(P, g["propose_phi"], g["c0"], g["mu0"], float(g["dt_probe"]), float(g["sc0"][9]))
    # bitwise against the oracle in the same reduction order
    assert (d["ok"], d["it"]) == (o["ok"], o["it"])
    assert d["residual"] == o["residual"] and d["closure"] == o["closure"]
    assert np.array_equal(d["carbon"], o["carbon"]) and np.array_equal(d["mu"], o["mu"])
    # against the reference's sequential order
    assert (d["ok"], d["it"]) == (int(g["diff_ok"]), int(g["diff_it"]))
    if d["ok"]:
        assert np.abs(d["mu"] - g["diff_mu"]).max() <= MU_TOL
        assert np.abs(d["carbon"] - g["diff_carbon"]).max() <= 1e-14


# (fixture, cg_rtol, k): with cg_atol = 0 and cg_rtol near the rounding floor
# the recursive residual passes the tolerance while the true residual does
# not (:1053-1065): the solve restarts from the true residual once and then
# converges, or (the last case) restarts until cg_max and fails.  Counted
# with an instrumented copy of the oracle: one failed check in each of the
# first five cases, 997 in the last.
RESTART_CASES = [("small_48x40", 1e-16, 1), ("small_48x40", 1e-16, 4), ("colony_256x128", 1e-16, 4),
                 ("ragged_257x9", 2e-16, 4), ("window_alpha_32x40", 1e-16, 1), ("colony_256x128", 5e-17, 4)]


@pytest.mark.parametrize("name,rtol,k", RESTART_CASES)
def test_true_residual_check_failure(pkg, O, name, rtol, k):
    """The true-residual check that finalize carries: where it fails, the
    carbon update it formed is dropped, the true-residual pass recomputes r
    and z and the solve restarts.  Result, iteration count and residual
    bitwise the oracle's in the same blocked order; the separate pass ran."""
    g = load_golden(name)
    pd = dict(g["pd"])
    pd["cg_rtol"], pd["cg_atol"] = rtol, 0.0
    say("true residual check failure", name, rtol, "k", k)
    with pkg.Solver(pd, float(g["sc0"][9]), cells_per_thread=k) as s:
        s.set_state(g["phi0"], g["c0"], g["mu0"], g["sc0"][:9])
        d = s.implicit_diffusion(g["propose_phi"], float(g["dt_probe"]))
        passes = s.kernel_stats()["true_residual"]["launches"]
    P = O.params_from_dict(pd, red_mode=O.RED_BLOCKED, red_k=k)
    o = O.implicit_diffusion(P, g["propose_phi"], g["c0"], g["mu0"], float(g["dt_probe"]), float(g["sc0"][9]))
    assert (d["ok"], d["it"]) == (o["ok"], o["it"])
    assert d["residual"] == o["residual"] and d["closure"] == o["closure"]
    assert passes >= 1
    if o["ok"]:
        assert np.array_equal(d["carbon"], o["carbon"]) and np.array_equal(d["mu"], o["mu"])
    else:
        assert o["it"] == int(pd["cg_max"])


@pytest.mark.parametrize("name,atol,iters", [("small_48x40", 1e30, 0), ("small_48x40", 1e-4, 1),
                                             ("small_48x40", 1e-6, 2), ("colony_256x128", 1e30, 0),
                                             ("colony_256x128", 1e-4, 1), ("colony_256x128", 1e-6, 2)])
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_short_solves(pkg, O, name, atol, iters, precision):
    """Solves of 0, 1 and 2 PCG iterations (cg_atol large; closure_tol 1 so
    the short solve is accepted): finalize reads the solution as the last
    update left it -- mu_old - mu_ref before the x plane is first written,
    with or without the last step alpha p still to apply.  FP64 bitwise the
    oracle in the same blocked order, with the iteration counts the oracle
    takes; the mixed form takes the same number of iterations, and its
    iterate differs from the FP64 one by the FP32 rounding of its directions
    (2^-24 relative) times the steps taken: bar 1e-6 of the solve's largest
    change of mu (exactly equal after zero iterations)."""
    g = load_golden(name)
    pd = dict(g["pd"])
    pd["cg_atol"], pd["closure_tol"] = atol, 1.0
    say("short solve", name, atol, iters, precision)
    prec = pkg.FP64 if precision == "fp64" else pkg.MIXED
    with pkg.Solver(pd, float(g["sc0"][9]), precision=prec) as s:
        s.set_state(g["phi0"], g["c0"], g["mu0"], g["sc0"][:9])
        k = s.layout()[1]
        d = s.implicit_diffusion(g["propose_phi"], float(g["dt_probe"]))
    P = O.params_from_dict(pd, red_mode=O.RED_BLOCKED, red_k=k)
    o = O.implicit_diffusion(P, g["propose_phi"], g["c0"], g["mu0"], float(g["dt_probe"]), float(g["sc0"][9]))
    assert o["it"] == iters and o["ok"] == 1
    assert (d["ok"], d["it"]) == (o["ok"], o["it"])
    if precision == "fp64":
        assert d["residual"] == o["residual"] and d["closure"] == o["closure"]
        assert np.array_equal(d["carbon"], o["carbon"]) and np.array_equal(d["mu"], o["mu"])
    else:
        moved = np.abs(o["mu"] - g["mu0"]).max()
        dmu = np.abs(d["mu"] - o["mu"]).max()
        say("  mixed: max|dmu|", dmu, "of a change of", moved)
        assert dmu <= 1e-6 * moved if iters else dmu == 0.0


@pytest.mark.parametrize("cg_max", [2000, 9, 8])
def test_true_residual_check_failure_in_steps(pkg, O, cg_max):
    """Whole accepted steps whose solves restart after a failed true-residual
    check (cg_rtol 1e-16, cg_atol 0); with cg_max 9 and 8 some attempts end
    unconverged, a restart on their last allowed iteration among them, and
    are retried with half the step: step infos, fields and scalars bitwise the
    oracle's."""
    g = load_golden("small_48x40")
    pd = dict(g["pd"])
    pd["cg_rtol"], pd["cg_atol"], pd["cg_max"] = 1e-16, 0.0, cg_max
    end, steps = float(g["end_time"]), 4
    say("true residual check failure in steps, cg_max", cg_max)
    with pkg.Solver(pd, float(g["sc0"][9])) as s:
        s.set_state(g["phi0"], g["c0"], g["mu0"], g["sc0"][:9])
        k = s.layout()[1]
        infos = [s.advance(end) for _ in range(steps)]
        phi, c, mu, sc = s.get_state()
        passes = s.kernel_stats()["true_residual"]["launches"]
    P = O.params_from_dict(pd, red_mode=O.RED_BLOCKED, red_k=k)
    S = O.OracleState(g["phi0"], g["c0"], g["mu0"], g["sc0"][:9])
    for kk in range(steps):
        rc, oi = O.advance(P, S, float(g["sc0"][9]), end)
        assert rc == 0
        gi = infos[kk]
        assert [gi.accepted_dt, gi.linear_residual, gi.closure, gi.cg_iterations, gi.retries] == \
            [oi.accepted_dt, oi.linear_residual, oi.closure, oi.cg_iterations, oi.retries], kk
    assert passes >= 1
    assert np.array_equal(phi, S.phi) and np.array_equal(c, S.carbon) and np.array_equal(mu, S.mu)
    assert sc.to_list() == S.scalars()


@pytest.mark.parametrize("name", CASES)
def test_state_reductions(pkg, O, name):
    g = load_golden(name)
    say("state_reductions", name)
    with make(pkg, g) as s:
        k = s.layout()[1]
        e, m, v = s.reduced_energy(), s.total_carbon(), s.valid_state()
    P = blocked_params(O, g, k)
    eo, st = O.reduced_energy(P, g["phi0"], g["c0"], float(g["sc0"][9]))
    assert e == eo
    assert m == O.total_carbon(P, g["c0"])
    assert v == bool(g["valid0"])
    # sequential vs blocked summation: bounded by n*eps relative (n <= 32768);
    # the reference's own mass test allows 1e-11 relative (pearl_serial.cpp:1152)
    assert abs(e - float(g["energy0"])) <= 1e-12 * abs(float(g["energy0"]))
    assert abs(m - float(g["carbon0"])) <= 1e-12 * abs(float(g["carbon0"]))


@pytest.mark.parametrize("name", CASES)
def test_trajectory(pkg, O, name):
    g = load_golden(name)
    steps, end = int(g["steps"]), float(g["end_time"])
    say("trajectory", name, steps)
    with make(pkg, g) as s:
        k = s.layout()[1]
        infos = [s.advance(end) for _ in range(steps)]
        phi, c, mu, sc = s.get_state()
    # bitwise against the oracle in the blocked order
    P = blocked_params(O, g, k)
    S = O.OracleState(g["phi0"], g["c0"], g["mu0"], g["sc0"][:9])
    for kk in range(steps):
        rc, oi = O.advance(P, S, float(g["sc0"][9]), end)
        assert rc == 0
        gi = infos[kk]
        assert [gi.accepted_dt, gi.max_phase_change, gi.linear_residual, gi.closure, gi.energy_change,
                gi.cg_iterations, gi.retries] == [oi.accepted_dt, oi.max_phase_change, oi.linear_residual,
                                                  oi.closure, oi.energy_change, oi.cg_iterations, oi.retries], kk
    assert np.array_equal(phi, S.phi) and np.array_equal(c, S.carbon) and np.array_equal(mu, S.mu)
    assert sc.to_list() == S.scalars()
    # against the reference
    ref = g["step_infos"]
    assert [i.retries for i in infos] == [int(r[6]) for r in ref]
    assert [i.cg_iterations for i in infos] == [int(r[5]) for r in ref]
    assert [i.accepted_dt for i in infos] == [r[0] for r in ref]
    assert np.abs(phi - g["phi1"]).max() <= PHI_TOL
    assert np.abs(mu - g["mu1"]).max() <= MU_TOL
    assert sc.to_list()[6:] == list(g["sc1"][6:9])


@pytest.mark.parametrize("name", ["colony_256x128", "ragged_257x9", "tiny_3x2"])
@pytest.mark.parametrize("k", [8, 16])
def test_trajectory_tall_tiles(pkg, O, name, k):
    """Whole steps with tiles of 8 and 16 rows (the LDS tiles of prepare,
    cg_apply, the true residual and finalize hold k + 2 rows): bitwise equal
    to the oracle in the same blocked order."""
    g = load_golden(name)
    steps, end = int(g["steps"]), float(g["end_time"])
    say("trajectory", name, "k", k)
    with make(pkg, g, k) as s:
        infos = [s.advance(end) for _ in range(steps)]
        phi, c, mu, sc = s.get_state()
    P = blocked_params(O, g, k)
    S = O.OracleState(g["phi0"], g["c0"], g["mu0"], g["sc0"][:9])
    for kk in range(steps):
        rc, oi = O.advance(P, S, float(g["sc0"][9]), end)
        assert rc == 0
        assert infos[kk].cg_iterations == oi.cg_iterations and infos[kk].accepted_dt == oi.accepted_dt, kk
    assert np.array_equal(phi, S.phi) and np.array_equal(c, S.carbon) and np.array_equal(mu, S.mu)
    assert sc.to_list() == S.scalars()


def mirrored_colony(g, rx, ry):
    """The colony fixture repeated rx times in x and ry times in y, every
    second copy flipped in y, so each seam joins equal rows and the state
    stays a valid, smooth colony."""
    def big(a):
        col = np.concatenate([a if i % 2 == 0 else a[..., ::-1, :] for i in range(ry)], axis=-2)
        return np.ascontiguousarray(np.tile(col, (1,) * (a.ndim - 1) + (rx,)))
    return big(g["phi0"]), big(g["c0"]), big(g["mu0"])


@pytest.mark.parametrize("phase_only", [0, 1])
def test_many_tiles_bitwise(pkg, O, phase_only):
    """2048x1152 cells in one-row tiles: 9,216 tiles, more than one round of
    a finisher's partial fetch (8,192 partials per round for one or two
    quantities, 4,096 for three or four, 2,048 for five), so every fold of
    the blocked reductions runs several rounds.  Two whole steps (coupled:
    propose, the PCG and the acceptance sums; phase-only: the five-quantity
    acceptance fold) and the direct reductions are bitwise equal to the
    oracle in the same blocked order."""
    g = load_golden("colony_256x128")
    rx, ry, k = 8, 9, 1
    phi0, c0, mu0 = mirrored_colony(g, rx, ry)
    pd = dict(g["pd"])
    pd["nx"], pd["ny"], pd["phase_only"] = 256 * rx, 128 * ry, phase_only
    mu_ref, end = float(g["sc0"][9]), float(g["end_time"])
    say("many tiles", pd["nx"], pd["ny"], "k", k, "phase_only", phase_only)
    with pkg.Solver(pd, mu_ref, cells_per_thread=k) as s:
        s.set_state(phi0, c0, mu0, g["sc0"][:9])
        assert s.layout()[1] == k and (pd["nx"] // 256) * pd["ny"] > 8192
        m = s.total_carbon()
        e = None if phase_only else s.reduced_energy()
        infos = [s.advance(end) for _ in range(2)]
        phi, c, mu, sc = s.get_state()
    P = O.params_from_dict(pd, red_mode=O.RED_BLOCKED, red_k=k)
    assert m == O.total_carbon(P, c0)
    if not phase_only:
        assert e == O.reduced_energy(P, phi0, c0, mu_ref)[0]
    S = O.OracleState(phi0, c0, mu0, g["sc0"][:9])
    for kk in range(2):
        rc, oi = O.advance(P, S, mu_ref, end)
        assert rc == 0
        gi = infos[kk]
        assert [gi.accepted_dt, gi.max_phase_change, gi.linear_residual, gi.closure, gi.energy_change,
                gi.cg_iterations, gi.retries] == [oi.accepted_dt, oi.max_phase_change, oi.linear_residual,
                                                  oi.closure, oi.energy_change, oi.cg_iterations, oi.retries], kk
    assert np.array_equal(phi, S.phi) and np.array_equal(c, S.carbon) and np.array_equal(mu, S.mu)
    assert sc.to_list() == S.scalars()


@pytest.mark.parametrize("name", ["window_32x40", "window_alpha_32x40"])
def test_moving_window(pkg, name):
    g = load_golden(name)
    say("moving_window", name)
    with pkg.Solver(g["pd"], float(g["sc0"][9])) as s:
        s.set_state(g["phi1"], g["c1"], g["mu1"], g["sc1"][:9])
        err = str(g["shift_error"])
        if err:
            with pytest.raises(pkg.PearlError) as ei:
                s.maybe_shift(True, *[float(v) for v in g["window"]])
            assert str(ei.value) == err
            return
        shifted, front = s.maybe_shift(True, *[float(v) for v in g["window"]])
        phi, c, mu, sc = s.get_state()
    assert shifted
    assert np.array_equal(phi, g["phi2"]) and np.array_equal(c, g["c2"])
    assert np.array_equal(mu, g["mu2"])
    ref = list(g["sc2"][:9])
    got = sc.to_list()
    assert got[:2] == ref[:2] and got[2] == ref[2] and got[4] == ref[4] and got[6:] == ref[6:]
    assert abs(got[5] - ref[5]) <= 1e-12 * abs(ref[5])


def test_shift_after_failed_shift(pkg):
    """A shift that fails inside the row move (a non-finite carbon value
    moved down: "Nonfinite total carbon") leaves no flag behind: a later
    shift that removes that cell instead succeeds and equals the same shift
    on a fresh context."""
    g = load_golden("window_alpha_32x40")
    trigger, target, c_gamma = [float(v) for v in g["window"]]
    dy = float(g["pd"]["dy"])
    nx = int(g["pd"]["nx"])
    c1 = np.array(g["c1"], dtype=np.float64).reshape(-1)
    c1[20 * nx + 5] = np.nan  # row 20: moved by the 18-row shift, removed by a 22-row one
    say("shift after a failed shift")
    with pkg.Solver(g["pd"], float(g["sc0"][9])) as s:
        s.set_state(g["phi1"], c1, g["mu1"], g["sc1"][:9])
        with pytest.raises(pkg.PearlError) as ei:
            s.maybe_shift(True, trigger, target, c_gamma)
        assert str(ei.value) == "Nonfinite total carbon"
        shifted, _ = s.maybe_shift(True, trigger, target - 4 * dy, c_gamma)
        assert shifted
        got = s.get_state()
    with pkg.Solver(g["pd"], float(g["sc0"][9])) as s:
        s.set_state(g["phi1"], c1, g["mu1"], g["sc1"][:9])
        assert s.maybe_shift(True, trigger, target - 4 * dy, c_gamma)[0]
        want = s.get_state()
    for a, b in zip(got[:3], want[:3]):
        assert np.all(np.isfinite(a)) and np.array_equal(a, b)
    assert got[3].shifts == want[3].shifts == 1


def test_diagnostics(pkg, O):
    g = load_golden("small_48x40")
    say("diagnostics")
    with make(pkg, g) as s:
        d = s.diagnostics()
        k = s.layout()[1]
    P = blocked_params(O, g, k)
    assert d["c_min"] == g["c0"].min() and d["c_max"] == g["c0"].max()
    assert d["mu_min"] == g["mu0"].min() and d["mu_max"] == g["mu0"].max()
    ph = g["phi0"]
    assert d["phi_sum_error"] == np.abs(ph[0] + ph[1] + ph[2] - 1).max()
    assert d["mass"] == O.total_carbon(P, g["c0"])
    assert d["energy"] == O.reduced_energy(P, g["phi0"], g["c0"], float(g["sc0"][9]))[0]


def test_forced_rejection_paths(pkg):
    """Host-forced failures: a tiny max_phase_change rejects every attempt
    on the phase increment until dt falls below dt_min; cg_max=0 rejects on
    linear/closure (the reference's messages, pearl_serial.cpp:1171)."""
    g = load_golden("small_48x40")
    say("forced_rejections")
    pd = dict(g["pd"])
    pd["max_phase_change"] = 1e-12
    with pkg.Solver(pd, float(g["sc0"][9])) as s:
        s.set_state(g["phi0"], g["c0"], g["mu0"], g["sc0"][:9])
        with pytest.raises(pkg.PearlError) as ei:
            s.advance(float(g["end_time"]))
        assert ei.value.code == pkg.ERR_DT_MIN
        assert str(ei.value) == "step rejected below dt_min: phase increment"
    pd = dict(g["pd"])
    pd["cg_max"] = 0
    with pkg.Solver(pd, float(g["sc0"][9])) as s:
        s.set_state(g["phi0"], g["c0"], g["mu0"], g["sc0"][:9])
        with pytest.raises(pkg.PearlError) as ei:
            s.advance(float(g["end_time"]))
        assert str(ei.value) in ("step rejected below dt_min: linear/closure", "step retry limit: linear/closure")
    with pkg.Solver(g["pd"], float(g["sc0"][9])) as s:
        s.set_state(g["phi0"], g["c0"], g["mu0"], g["sc0"][:9])
        with pytest.raises(pkg.PearlError) as ei:
            s.advance(0.0)
        assert str(ei.value) == "invalid advance time"


def test_full_smoke_trajectory_vs_reference(pkg):
    """The 256x128 lamellar smoke case to 0.015 us (the reference's own
    horizon), from the reference's initialized state: identical accepted
    steps, rejected attempts and PCG iteration total, and final fields
    within the FP64 bar of the reference run."""
    g = load_golden("colony_256x128_full")
    end = float(g["end_time"])
    say("full trajectory 256x128 to", end)
    with pkg.Solver(g["pd"], float(g["sc0"][9])) as s:
        s.set_state(g["phi0"], g["c0"], g["mu0"], g["sc0"][:9])
        r = s.run(end, 10 ** 9)
        phi, c, mu, sc = s.get_state()
    say("  steps", r["taken"], "cg", r["cg_total"], "rejected", sc.rejected)
    assert r["taken"] == int(g["taken"])
    assert r["cg_total"] == int(g["cg_total"]) and r["cg_peak"] == int(g["cg_peak"])
    assert sc.rejected == int(g["sc1"][8]) and sc.step == int(g["sc1"][6])
    assert abs(sc.time - float(g["sc1"][0])) <= 1e-12 * end
    dphi = np.abs(phi - g["phi1"]).max()
    dmu = np.abs(mu - g["mu1"]).max()
    say("  max|dphi|", dphi, "max|dmu|", dmu, "max|dc|", np.abs(c - g["c1"]).max())
    assert dphi <= PHI_TOL
    assert dmu <= MU_TOL


@pytest.mark.parametrize("names", [["small_48x40", "window_alpha_32x40", "colony_256x128"],
                                   ["small_48x40", "phase_only_48x40", "colony_256x128"]])
def test_batch_of_uneven_domains_matches_single_domains(pkg, O, names):
    """Three domains of different sizes and physics in one context: each is
    bitwise equal to its own oracle trajectory in the context's blocked
    order.  With a phase-only domain in the context, the round ends in the
    phase-only pass, whose finisher takes every domain's turn (the coupled
    domains' too)."""
    gs = [load_golden(n) for n in names]
    steps = [12, 5, 8]
    say("batch of", names)
    with pkg.Batch([g["pd"] for g in gs], [float(g["sc0"][9]) for g in gs]) as b:
        k = b.layout()[1]
        for d, g in enumerate(gs):
            b.set_domain_state(d, g["phi0"], g["c0"], g["mu0"], g["sc0"][:9])
        r = b.run_all([float(g["end_time"]) for g in gs], steps)
        assert r["rc"] == 0, r
        assert r["taken"] == steps
        states = [b.get_domain_state(d) for d in range(len(gs))]
    for d, g in enumerate(gs):
        P = blocked_params(O, g, k)
        S = O.OracleState(g["phi0"], g["c0"], g["mu0"], g["sc0"][:9])
        for _ in range(steps[d]):
            rc, _ = O.advance(P, S, float(g["sc0"][9]), float(g["end_time"]))
            assert rc == 0
        phi, c, mu, sc = states[d]
        assert np.array_equal(phi, S.phi) and np.array_equal(c, S.carbon) and np.array_equal(mu, S.mu), names[d]
        assert sc.to_list() == S.scalars(), names[d]


def test_profiled_and_graph_runs_agree(pkg):
    """Eager (profiled) rounds and captured-graph rounds give the same bits."""
    g = load_golden("small_48x40")
    out = []
    for prof in (False, True):
        with make(pkg, g) as s:
            s.profile(prof)
            s.run(float(g["end_time"]), 15)
            out.append(s.get_state())
            if prof:
                st = s.kernel_stats()
                assert st["propose_phi"]["launches"] >= 15 and st["cg_apply"]["bytes"] > 0
    for a, b in zip(out[0][:3], out[1][:3]):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_batch_with_restarting_domain(pkg, O, precision):
    """Domains that share one set of launches while their solves take
    different paths: the first domain's solves fail the true-residual check
    inside finalize and restart (cg_rtol 1e-16, cg_atol 0), the second solves
    normally, the third converges in one iteration (cg_atol 1e-4,
    closure_tol 1).  Each domain bitwise the same domain alone in its own
    context; the FP64 batch also bitwise the oracle in the blocked order."""
    names = ["small_48x40", "colony_256x128", "window_alpha_32x40"]
    gs = [load_golden(n) for n in names]
    pds = [dict(g["pd"]) for g in gs]
    pds[0]["cg_rtol"], pds[0]["cg_atol"] = 1e-16, 0.0
    pds[2]["cg_atol"], pds[2]["closure_tol"] = 1e-4, 1.0
    steps = [4, 6, 3]
    prec = pkg.FP64 if precision == "fp64" else pkg.MIXED
    say("batch with a restarting domain", precision)
    with pkg.Batch(pds, [float(g["sc0"][9]) for g in gs], precision=prec) as b:
        k = b.layout()[1]
        for d, g in enumerate(gs):
            b.set_domain_state(d, g["phi0"], g["c0"], g["mu0"], g["sc0"][:9])
        r = b.run_all([float(g["end_time"]) for g in gs], steps)
        assert r["rc"] == 0, r
        assert r["taken"] == steps
        passes = b.kernel_stats()["true_residual"]["launches"]
        states = [b.get_domain_state(d) for d in range(len(gs))]
    assert passes >= steps[0]
    for d, g in enumerate(gs):
        with pkg.Solver(pds[d], float(g["sc0"][9]), cells_per_thread=k, precision=prec) as s:
            s.set_state(g["phi0"], g["c0"], g["mu0"], g["sc0"][:9])
            for _ in range(steps[d]):
                s.advance(float(g["end_time"]))
            alone = s.get_state()
        for a, bb in zip(states[d][:3], alone[:3]):
            assert np.array_equal(a, bb), names[d]
        assert states[d][3].to_list() == alone[3].to_list(), names[d]
        if precision == "fp64":
            P = O.params_from_dict(pds[d], red_mode=O.RED_BLOCKED, red_k=k)
            S = O.OracleState(g["phi0"], g["c0"], g["mu0"], g["sc0"][:9])
            for _ in range(steps[d]):
                rc, _ = O.advance(P, S, float(g["sc0"][9]), float(g["end_time"]))
                assert rc == 0
            phi, c, mu, sc = states[d]
            assert np.array_equal(phi, S.phi) and np.array_equal(c, S.carbon) and np.array_equal(mu, S.mu), names[d]
            assert sc.to_list() == S.scalars(), names[d]
