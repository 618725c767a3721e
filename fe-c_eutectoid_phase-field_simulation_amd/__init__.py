"""Pearl per-step solver on MI355X: Python binding of the C ABI.

The host-side mirror of the reference's step interface (pearl_serial.cpp):

  Solver.advance(stop_time)          advance            :1119-1174
  Solver.maybe_shift(...)            maybe_shift        :2053-2071
  Solver.propose_phi(dt)             propose_phi        :961-1002
  Solver.implicit_diffusion(phi, dt) implicit_diffusion :1010-1094
  Solver.reduced_energy()            reduced_energy     :793-821
  Solver.total_carbon()              total_carbon       :784-788
  Solver.valid_state()               valid_state        :1096-1116
  Solver.diagnostics()               snapshot reductions :2267-2282

Errors raise PearlError carrying the reference's exception text.  Every call
runs the HIP kernels of libpearl_gpu.so; importing the package on a machine
without the built library raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG_DIR, "libpearl_gpu.so")

OK = 0
ERR_INVALID_TIME, ERR_DT_MIN, ERR_RETRY_LIMIT, ERR_MIXTURE, ERR_PROJECTION = 1, 2, 3, 4, 5
ERR_UNSAFE_SHIFT, ERR_NONFINITE_CARBON, ERR_INVALID_SHIFT, ERR_ALLOC, ERR_ARGUMENT, ERR_DEVICE = 6, 7, 8, 9, 10, 11
WHY = ["phase increment", "linear/closure", "composition/phase bounds", "mass conservation",
       "free-energy increase"]
FP64 = 0
MIXED = 1  # Jacobi diagonal, p, z in FP32; state, operator, residuals, dots and acceptance FP64
FP32_STENCIL = 2  # MIXED plus FP32 arithmetic in the phase proposal's stencil (force, mean, step scale)

KERNELS = ["propose_phi", "diffusion_prepare", "cg_apply", "cg_update", "true_residual",
           "idle", "finalize_accept", "state_reduce", "moving_window", "reduce_finish", "cg_apply_first"]

_dp = C.POINTER(C.c_double)


class PearlError(RuntimeError):
    def __init__(self, code: int, message: str):
        super().__init__(message)
        self.code = code


class Params(C.Structure):
    """pearl_params (include/pearl_gpu.h)."""
    _fields_ = [
        ("nx", C.c_int), ("ny", C.c_int), ("dx", C.c_double), ("dy", C.c_double),
        ("A", C.c_double * 3), ("B", C.c_double * 3), ("C", C.c_double * 3), ("D", C.c_double * 3),
        ("eps", C.c_double), ("sigma", C.c_double), ("triple", C.c_double),
        ("tau", C.c_double * 3), ("Vm", C.c_double),
        ("dt_max", C.c_double), ("dt_min", C.c_double), ("max_phase_change", C.c_double),
        ("cg_rtol", C.c_double), ("cg_atol", C.c_double), ("closure_tol", C.c_double),
        ("energy_rtol", C.c_double), ("energy_atol", C.c_double),
        ("cg_max", C.c_int), ("max_retries", C.c_int), ("check_energy", C.c_int), ("phase_only", C.c_int),
        ("pair_sigma", C.c_double * 3),
    ]


class Scalars(C.Structure):
    _fields_ = [
        ("time", C.c_double), ("dt", C.c_double), ("shift_distance", C.c_double),
        ("initial_carbon", C.c_double), ("carbon_added", C.c_double), ("carbon_removed", C.c_double),
        ("step", C.c_longlong), ("shifts", C.c_longlong), ("rejected", C.c_longlong),
    ]
    NAMES = ["time", "dt", "shift_distance", "initial_carbon", "carbon_added", "carbon_removed",
             "step", "shifts", "rejected"]

    @classmethod
    def from_list(cls, v):
        s = cls()
        for k, x in zip(cls.NAMES, v):
            setattr(s, k, int(x) if k in ("step", "shifts", "rejected") else float(x))
        return s

    def to_list(self):
        return [float(getattr(self, k)) for k in self.NAMES]


class StepInfo(C.Structure):
    _fields_ = [
        ("accepted_dt", C.c_double), ("max_phase_change", C.c_double), ("linear_residual", C.c_double),
        ("closure", C.c_double), ("energy_change", C.c_double),
        ("cg_iterations", C.c_int), ("retries", C.c_int), ("why", C.c_int),
    ]


class DiffusionInfo(C.Structure):
    _fields_ = [("ok", C.c_int), ("it", C.c_int), ("residual", C.c_double), ("closure", C.c_double)]


class Diag(C.Structure):
    _fields_ = [("c_min", C.c_double), ("c_max", C.c_double), ("mu_min", C.c_double), ("mu_max", C.c_double),
                ("phi_sum_error", C.c_double), ("mass", C.c_double), ("energy", C.c_double)]


_lib = None

# Every symbol declared in include/pearl_gpu.h.
class SmoothReport(C.Structure):
    _fields_ = [("steps", C.c_int), ("converged", C.c_int), ("energy_initial", C.c_double),
                ("energy_final", C.c_double), ("final_relative_energy_change", C.c_double),
                ("final_max_phase_change", C.c_double)]


class Fill(C.Structure):
    """pearl_fill: one FILL operation (filling_phase_at, pearl_serial.cpp:1393-1514)."""
    _fields_ = [("method", C.c_int), ("reserved", C.c_int), ("p", C.c_double * 16),
                ("dir", (C.c_double * 2) * 3)]


# Method codes of pearl_fill, in the order of the PEARL_FILL_* constants.
FILL_METHODS = ("FILLCUBE", "FILLCYLINDER", "FILLCENTERBOX2D", "FILLYJUNCTION2D", "FILLYJUNCTIONLAMELLAE2D",
                "FILLADHEREDPEARLITEARM2D", "FILLPEARLITEELLIPSE2D")
COMPOSITION_MU, COMPOSITION_MEAN = 0, 1
PROFILE_LAMELLA, PROFILE_FLAT, PROFILE_CELLS = 0, 1, 2


def fill_ops(cfg):
    """The sharp labelling of `cfg` (a config.Config read from its input
    file) as pearl_fill operations, formed by the input reader of the
    drop-in program (libpearl_input.so): the admitted FILL lines for
    initialization = microsim, with each direction the reference's
    cos(angle*pi/180.0) (:1396), or the two boxes of sharp_smooth (theta in
    the first theta_columns, alpha in the rest, of the rows whose centre
    (gy + 0.5) dy lies below seed_height, pearl_serial.cpp:1777-1786).
    Returns (array of Fill, count)."""
    from . import config
    try:
        return config.sharp_label_ops(cfg, Fill)
    except config.ConfigError as e:
        raise ValueError(str(e)) from None


EXPORTS = ["pearl_create", "pearl_destroy", "pearl_last_error", "pearl_set_state", "pearl_get_state",
           "pearl_advance", "pearl_run", "pearl_maybe_shift", "pearl_diagnostics", "pearl_propose_phi",
           "pearl_implicit_diffusion", "pearl_reduced_energy", "pearl_total_carbon", "pearl_valid_state",
           "pearl_get_layout", "pearl_profile", "pearl_kernel_stats", "pearl_kernel_name", "pearl_counters",
           "pearl_batch_create", "pearl_batch_size", "pearl_batch_set_state", "pearl_batch_get_state",
           "pearl_batch_run", "pearl_smooth", "pearl_step_fixed", "pearl_adjoint", "pearl_set_stop_tolerance",
           "pearl_run_fixed", "pearl_fill_phases", "pearl_init_composition", "pearl_stream_probe",
           "pearl_diagnostics_noalloc", "pearl_init_profile", "pearl_stream_mix"]


def lib():
    """Load libpearl_gpu.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
                               "(hipcc, gfx950); there is no CPU fallback")
        L = C.CDLL(LIB_PATH)
        vp = C.c_void_p
        L.pearl_create.argtypes = [C.POINTER(Params), C.c_double, C.c_int, C.c_int, C.c_int, C.POINTER(vp)]
        L.pearl_destroy.argtypes = [vp]
        L.pearl_last_error.restype = C.c_char_p
        L.pearl_last_error.argtypes = [vp]
        L.pearl_set_state.argtypes = [vp, _dp, _dp, _dp, C.POINTER(Scalars)]
        L.pearl_get_state.argtypes = [vp, _dp, _dp, _dp, C.POINTER(Scalars)]
        L.pearl_advance.argtypes = [vp, C.c_double, C.POINTER(StepInfo)]
        L.pearl_run.argtypes = [vp, C.c_double, C.c_longlong, C.POINTER(C.c_longlong), C.POINTER(C.c_longlong),
                                C.POINTER(C.c_int), C.POINTER(StepInfo)]
        L.pearl_maybe_shift.argtypes = [vp, C.c_int, C.c_double, C.c_double, C.c_double, C.POINTER(C.c_int), _dp]
        L.pearl_diagnostics.argtypes = [vp, C.POINTER(Diag)]
        L.pearl_diagnostics_noalloc.argtypes = [C.POINTER(Params), C.c_double, C.c_int, vp, vp, vp, vp,
                                                C.POINTER(Diag), C.c_char_p, C.c_int]
        L.pearl_smooth.argtypes = [vp, C.c_int, C.c_int, C.c_double, C.c_double, C.POINTER(SmoothReport)]
        L.pearl_fill_phases.argtypes = [vp, C.POINTER(Fill), C.c_int]
        L.pearl_init_profile.argtypes = [vp, C.c_int, _dp, _dp, _dp, C.c_int, C.c_int]
        L.pearl_stream_probe.argtypes = [C.c_int, C.c_longlong, C.c_int, _dp]
        L.pearl_stream_mix.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, _dp, _dp]
        L.pearl_init_composition.argtypes = [vp, C.c_int, C.c_double, _dp, _dp, C.POINTER(C.c_int)]
        L.pearl_step_fixed.argtypes = [vp, C.c_double]
        L.pearl_set_stop_tolerance.argtypes = [vp, C.c_double]
        L.pearl_run_fixed.argtypes = [vp, C.c_int, _dp]
        L.pearl_adjoint.argtypes = [vp, C.c_int, _dp, C.c_int, _dp, C.c_double, C.c_double, _dp, _dp]
        L.pearl_propose_phi.argtypes = [vp, C.c_double, _dp, _dp]
        L.pearl_implicit_diffusion.argtypes = [vp, _dp, C.c_double, _dp, _dp, C.POINTER(DiffusionInfo)]
        L.pearl_reduced_energy.argtypes = [vp, _dp]
        L.pearl_total_carbon.argtypes = [vp, _dp]
        L.pearl_valid_state.argtypes = [vp, C.POINTER(C.c_int)]
        L.pearl_get_layout.argtypes = [vp, C.POINTER(C.c_int), C.POINTER(C.c_int)]
        L.pearl_profile.argtypes = [vp, C.c_int]
        L.pearl_kernel_stats.argtypes = [vp, C.c_int, C.POINTER(C.c_longlong), _dp, _dp]
        L.pearl_kernel_name.restype = C.c_char_p
        L.pearl_kernel_name.argtypes = [C.c_int]
        L.pearl_counters.argtypes = [vp, C.POINTER(C.c_longlong), C.POINTER(C.c_longlong), C.POINTER(C.c_longlong)]
        L.pearl_batch_create.argtypes = [C.c_int, C.POINTER(Params), _dp, C.c_int, C.c_int, C.c_int, C.POINTER(vp)]
        L.pearl_batch_size.argtypes = [vp, C.POINTER(C.c_int)]
        L.pearl_batch_set_state.argtypes = [vp, C.c_int, _dp, _dp, _dp, C.POINTER(Scalars)]
        L.pearl_batch_get_state.argtypes = [vp, C.c_int, _dp, _dp, _dp, C.POINTER(Scalars)]
        L.pearl_batch_run.argtypes = [vp, _dp, C.POINTER(C.c_longlong), C.POINTER(C.c_longlong),
                                      C.POINTER(C.c_longlong), C.POINTER(C.c_int), C.POINTER(StepInfo),
                                      C.POINTER(C.c_int)]
        _lib = L
    return _lib


PARAM_KEYS = ["A0", "B0", "C0", "D0", "A1", "B1", "C1", "D1", "A2", "B2", "C2", "D2",
              "eps", "sigma", "triple", "tau_at", "tau_ag", "tau_tg", "Vm", "dx", "dy", "nx", "ny",
              "dt_max", "dt_min", "max_phase_change", "cg_rtol", "cg_atol", "closure_tol",
              "energy_rtol", "energy_atol", "cg_max", "max_retries", "check_energy", "phase_only"]


PAIR_SIGMA_KEYS = ("sigma_at", "sigma_ag", "sigma_tg")
ADJOINT_GRAD = ("sigma", "tau_at", "tau_ag", "tau_tg", "D0", "D1", "D2", "sigma_at", "sigma_ag", "sigma_tg")


def stream_probe(device: int = 0, nbytes: int = 1 << 30, reps: int = 20) -> float:
    """GB/s of a plain device copy of `nbytes` (pearl_stream_probe: read +
    write counted), the measured bandwidth ceiling."""
    g = C.c_double(0)
    rc = lib().pearl_stream_probe(int(device), int(nbytes), int(reps), C.byref(g))
    if rc != OK:
        raise PearlError(rc, "pearl_stream_probe failed")
    return g.value


# FP64 planes each cell kernel reads and writes once per cell (the access
# pattern of its algorithmic bytes; the FP64 plain form).  cg_apply: the
# direction launches that also apply the previous step to x (all but the
# first of a solve, cg_apply_first); finalize_accept: with the last step
# applied; true_residual: the adjoint's pass.
KERNEL_STREAMS = {"propose_phi": (4, 3), "diffusion_prepare": (5, 5), "cg_apply": (5, 3), "cg_apply_first": (3, 1),
                  "cg_update": (3, 2), "true_residual": (5, 2), "finalize_accept": (7, 2)}


def stream_mix(reads: int, writes: int, nx: int, ny: int, rows_per_thread: int = 4, reps: int = 20,
               device: int = 0):
    """(us per launch, GB/s) of a bare kernel with `reads` and `writes` FP64
    planes in the cell kernels' geometry (pearl_stream_mix): the rate that
    access pattern reaches without arithmetic."""
    us, g = C.c_double(0), C.c_double(0)
    rc = lib().pearl_stream_mix(int(device), int(reads), int(writes), int(nx), int(ny), int(rows_per_thread),
                                int(reps), C.byref(us), C.byref(g))
    if rc != OK:
        raise PearlError(rc, "pearl_stream_mix failed")
    return us.value, g.value


def diagnostics_noalloc(params, mu_reference: float, phi_ptr: int, carbon_ptr: int, mu_ptr: int,
                        cells_per_thread: int = 0, stream: int = 0) -> dict:
    """pearl_diagnostics_noalloc: the snapshot reductions of a state in the
    caller's device memory (device addresses of phi[3,ny,nx], carbon, mu,
    float64, contiguous) on the current device; allocates no device memory."""
    P = params if isinstance(params, Params) else params_from_dict(params)
    d = Diag()
    msg = C.create_string_buffer(256)
    rc = lib().pearl_diagnostics_noalloc(C.byref(P), mu_reference, cells_per_thread, phi_ptr, carbon_ptr, mu_ptr,
                                         stream or None, C.byref(d), msg, 256)
    if rc != OK:
        raise PearlError(rc, msg.value.decode())
    return {k: getattr(d, k) for k, _ in Diag._fields_}


def params_from_dict(d: dict) -> Params:
    """Build pearl_params from a dict with the keys of PARAM_KEYS."""
    p = Params()
    p.nx, p.ny = int(d["nx"]), int(d["ny"])
    p.dx, p.dy = float(d["dx"]), float(d["dy"])
    for i in range(3):
        p.A[i], p.B[i], p.C[i], p.D[i] = d[f"A{i}"], d[f"B{i}"], d[f"C{i}"], d[f"D{i}"]
    p.eps, p.sigma, p.triple = d["eps"], d["sigma"], d["triple"]
    p.tau[0], p.tau[1], p.tau[2] = d["tau_at"], d["tau_ag"], d["tau_tg"]
    p.Vm = d["Vm"]
    for k in ("dt_max", "dt_min", "max_phase_change", "cg_rtol", "cg_atol", "closure_tol",
              "energy_rtol", "energy_atol"):
        setattr(p, k, float(d[k]))
    p.cg_max, p.max_retries = int(d["cg_max"]), int(d["max_retries"])
    p.check_energy, p.phase_only = int(d["check_energy"]), int(d["phase_only"])
    # optional extension: interface energy per phase pair (0 = sigma)
    for i, k in enumerate(PAIR_SIGMA_KEYS):
        p.pair_sigma[i] = float(d.get(k, 0.0))
    return p


def _ptr(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_dp)


class Solver:
    """One device-resident domain (a pearl_ctx)."""

    def __init__(self, params, mu_reference: float, device: int = 0, cells_per_thread: int = 0,
                 precision: int = FP64):
        if isinstance(params, dict):
            params = params_from_dict(params)
        self.params = params
        self.nx, self.ny = params.nx, params.ny
        self.mu_reference = float(mu_reference)
        h = C.c_void_p()
        rc = lib().pearl_create(C.byref(params), self.mu_reference, precision, device, cells_per_thread,
                                C.byref(h))
        if rc != OK:
            raise PearlError(rc, f"pearl_create failed with status {rc}")
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            lib().pearl_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _check(self, rc):
        if rc != OK:
            raise PearlError(rc, lib().pearl_last_error(self._h).decode())

    # state -----------------------------------------------------------------
    def set_state(self, phi, carbon, mu, scalars=None):
        phi = np.ascontiguousarray(phi, dtype=np.float64).reshape(3, self.ny, self.nx)
        carbon = np.ascontiguousarray(carbon, dtype=np.float64).reshape(self.ny, self.nx)
        mu = np.ascontiguousarray(mu, dtype=np.float64).reshape(self.ny, self.nx)
        sc = None
        if scalars is not None:
            sc = scalars if isinstance(scalars, Scalars) else Scalars.from_list(scalars)
        self._check(lib().pearl_set_state(self._h, _ptr(phi), _ptr(carbon), _ptr(mu),
                                          C.byref(sc) if sc is not None else None))

    def set_scalars(self, scalars):
        """Replace the State scalars only (null field pointers leave the
        fields and the cached e0/mass0 untouched), as the host driver does
        after the clock snap (pearl_serial.cpp:2309-2310)."""
        sc = scalars if isinstance(scalars, Scalars) else Scalars.from_list(scalars)
        self._check(lib().pearl_set_state(self._h, None, None, None, C.byref(sc)))

    def get_state(self):
        phi = np.empty((3, self.ny, self.nx))
        c = np.empty((self.ny, self.nx))
        mu = np.empty((self.ny, self.nx))
        sc = Scalars()
        self._check(lib().pearl_get_state(self._h, _ptr(phi), _ptr(c), _ptr(mu), C.byref(sc)))
        return phi, c, mu, sc

    def scalars(self) -> Scalars:
        sc = Scalars()
        self._check(lib().pearl_get_state(self._h, None, None, None, C.byref(sc)))
        return sc

    # the step -----------------------------------------------------------------
    def advance(self, stop_time: float) -> StepInfo:
        info = StepInfo()
        self._check(lib().pearl_advance(self._h, stop_time, C.byref(info)))
        return info

    def run(self, stop_time: float, max_steps: int):
        taken, cg_total, cg_peak = C.c_longlong(0), C.c_longlong(0), C.c_int(0)
        last = StepInfo()
        self._check(lib().pearl_run(self._h, stop_time, max_steps, C.byref(taken), C.byref(cg_total),
                                    C.byref(cg_peak), C.byref(last)))
        return dict(taken=taken.value, cg_total=cg_total.value, cg_peak=cg_peak.value, last=last)

    def set_stop_tolerance(self, eps: float):
        """run()/run_all() end once time >= stop_time - eps (pearl_set_stop_tolerance)."""
        self._check(lib().pearl_set_stop_tolerance(self._h, float(eps)))

    def maybe_shift(self, moving_window: bool, trigger: float, target: float, c_gamma: float):
        sh = C.c_int(0)
        fr = C.c_double(0)
        self._check(lib().pearl_maybe_shift(self._h, int(bool(moving_window)), trigger, target, c_gamma,
                                            C.byref(sh), C.byref(fr)))
        return bool(sh.value), fr.value

    def diagnostics(self) -> dict:
        d = Diag()
        self._check(lib().pearl_diagnostics(self._h, C.byref(d)))
        return {k: getattr(d, k) for k, _ in Diag._fields_}

    def step_fixed(self, dt: float):
        """One always-taken step of size dt (propose_phi, implicit_diffusion)."""
        self._check(lib().pearl_step_fixed(self._h, float(dt)))

    def run_fixed(self, dts):
        """len(dts) always-taken steps of the given sizes in one call
        (pearl_run_fixed); the same result as step_fixed per size."""
        dts = np.ascontiguousarray(dts, dtype=np.float64)
        self._check(lib().pearl_run_fixed(self._h, int(dts.size), _ptr(dts)))

    def adjoint(self, dts, phi_target, every: int = 100, w_front: float = 0.0, front_target: float = 0.0):
        """Loss and gradient (ADJOINT_GRAD order: sigma, tau_at, tau_ag,
        tau_tg, D_alpha, D_theta, D_gamma, sigma_at, sigma_ag, sigma_tg)
        after fixed steps dts from the current state; see pearl_adjoint.
        The context ends on the final state."""
        dts = np.ascontiguousarray(dts, dtype=np.float64)
        tgt = np.ascontiguousarray(phi_target, dtype=np.float64).reshape(3, self.ny, self.nx)
        loss = C.c_double(0)
        grad = np.zeros(len(ADJOINT_GRAD))
        self._check(lib().pearl_adjoint(self._h, int(dts.size), _ptr(dts), int(every), _ptr(tgt), float(w_front),
                                        float(front_target), C.byref(loss), _ptr(grad)))
        return loss.value, grad

    def smooth(self, max_steps: int = 100, min_steps: int = 100, energy_rtol: float = 1e-10,
               max_phase_change: float = 0.02) -> dict:
        """Capillarity-only smoothing of the current phases in place
        (smooth_interface_only, pearl_serial.cpp:869-958; defaults :1720-1723)."""
        r = SmoothReport()
        self._check(lib().pearl_smooth(self._h, int(max_steps), int(min_steps), float(energy_rtol),
                                       float(max_phase_change), C.byref(r)))
        return {k: getattr(r, k) for k, _ in SmoothReport._fields_}

    def fill_phases(self, ops, count: int | None = None):
        """Sharp FILL labels of the current phases (the labelling loop of
        initialize, pearl_serial.cpp:1764-1795); `ops` from fill_ops."""
        n = len(ops) if count is None else int(count)
        self._check(lib().pearl_fill_phases(self._h, ops, n))

    def init_profile(self, kind: int, col=None, row=None, cell=None, phase_a: int = 0, phase_b: int = 2):
        """Diffuse-profile phases on the device (initialize's lamella, flat
        and circle branches, pearl_serial.cpp:1767-1771, :1787-1791) from
        profile values: PROFILE_LAMELLA col[nx] and row[ny], PROFILE_FLAT
        row[ny], PROFILE_CELLS cell[ny, nx]."""
        def arr(a, size):
            if a is None:
                return None
            a = np.ascontiguousarray(a, dtype=np.float64).reshape(-1)
            if a.size != size:
                raise ValueError(f"profile array of {a.size} values, expected {size}")
            return a
        col = arr(col, self.nx)
        row = arr(row, self.ny)
        cell = arr(cell, self.nx * self.ny)
        self._check(lib().pearl_init_profile(self._h, int(kind), _ptr(col) if col is not None else None,
                                             _ptr(row) if row is not None else None,
                                             _ptr(cell) if cell is not None else None, int(phase_a), int(phase_b)))

    def init_composition(self, mode: int, value: float):
        """Uniform mu and the carbon of the current phases (initialize
        :1813-1836).  Returns (mu, initial_carbon, valid)."""
        mu, carbon, valid = C.c_double(0), C.c_double(0), C.c_int(0)
        self._check(lib().pearl_init_composition(self._h, int(mode), float(value), C.byref(mu), C.byref(carbon),
                                                 C.byref(valid)))
        return mu.value, carbon.value, bool(valid.value)

    # single reference functions ------------------------------------------------
    def propose_phi(self, dt: float):
        out = np.empty((3, self.ny, self.nx))
        ch = C.c_double(0)
        self._check(lib().pearl_propose_phi(self._h, dt, _ptr(out), C.byref(ch)))
        return ch.value, out

    def implicit_diffusion(self, phi_trial, dt: float):
        phi_trial = np.ascontiguousarray(phi_trial, dtype=np.float64).reshape(3, self.ny, self.nx)
        c = np.empty((self.ny, self.nx))
        mu = np.empty((self.ny, self.nx))
        info = DiffusionInfo()
        self._check(lib().pearl_implicit_diffusion(self._h, _ptr(phi_trial), dt, _ptr(c), _ptr(mu),
                                                   C.byref(info)))
        return dict(ok=info.ok, it=info.it, residual=info.residual, closure=info.closure, carbon=c, mu=mu)

    def reduced_energy(self) -> float:
        e = C.c_double(0)
        self._check(lib().pearl_reduced_energy(self._h, C.byref(e)))
        return e.value

    def total_carbon(self) -> float:
        m = C.c_double(0)
        self._check(lib().pearl_total_carbon(self._h, C.byref(m)))
        return m.value

    def valid_state(self) -> bool:
        v = C.c_int(0)
        self._check(lib().pearl_valid_state(self._h, C.byref(v)))
        return bool(v.value)

    # layout, timing ----------------------------------------------------------
    def layout(self):
        tpb, k = C.c_int(0), C.c_int(0)
        self._check(lib().pearl_get_layout(self._h, C.byref(tpb), C.byref(k)))
        return tpb.value, k.value

    def profile(self, enable: bool):
        self._check(lib().pearl_profile(self._h, int(bool(enable))))

    def kernel_stats(self) -> dict:
        out = {}
        for k, name in enumerate(KERNELS):
            n, ms, b = C.c_longlong(0), C.c_double(0), C.c_double(0)
            self._check(lib().pearl_kernel_stats(self._h, k, C.byref(n), C.byref(ms), C.byref(b)))
            out[name] = dict(launches=n.value, ms=ms.value, bytes=b.value)
        return out

    def counters(self):
        a, r, cg = C.c_longlong(0), C.c_longlong(0), C.c_longlong(0)
        self._check(lib().pearl_counters(self._h, C.byref(a), C.byref(r), C.byref(cg)))
        return dict(attempts=a.value, rejected_attempts=r.value, cg_iterations=cg.value)


class Batch(Solver):
    """Many independent domains of any sizes in one context (one packed HBM
    arena, one launch per kernel for all of them)."""

    def __init__(self, params_list, mu_references, device: int = 0, cells_per_thread: int = 0,
                 precision: int = FP64):
        ps = [p if isinstance(p, Params) else params_from_dict(p) for p in params_list]
        self.count = len(ps)
        arr = (Params * self.count)(*ps)
        mus = np.ascontiguousarray(mu_references, dtype=np.float64)
        h = C.c_void_p()
        rc = lib().pearl_batch_create(self.count, arr, _ptr(mus), precision, device, cells_per_thread, C.byref(h))
        if rc != OK:
            raise PearlError(rc, f"pearl_batch_create failed with status {rc}")
        self._h = h
        self.params = ps[0]
        self.plist = ps
        self.nx, self.ny = ps[0].nx, ps[0].ny

    def set_domain_state(self, d, phi, carbon, mu, scalars=None):
        p = self.plist[d]
        phi = np.ascontiguousarray(phi, dtype=np.float64).reshape(3, p.ny, p.nx)
        carbon = np.ascontiguousarray(carbon, dtype=np.float64).reshape(p.ny, p.nx)
        mu = np.ascontiguousarray(mu, dtype=np.float64).reshape(p.ny, p.nx)
        sc = None
        if scalars is not None:
            sc = scalars if isinstance(scalars, Scalars) else Scalars.from_list(scalars)
        self._check(lib().pearl_batch_set_state(self._h, d, _ptr(phi), _ptr(carbon), _ptr(mu),
                                                C.byref(sc) if sc is not None else None))

    def get_domain_state(self, d):
        p = self.plist[d]
        phi = np.empty((3, p.ny, p.nx))
        c = np.empty((p.ny, p.nx))
        mu = np.empty((p.ny, p.nx))
        sc = Scalars()
        self._check(lib().pearl_batch_get_state(self._h, d, _ptr(phi), _ptr(c), _ptr(mu), C.byref(sc)))
        return phi, c, mu, sc

    def run_all(self, stop_times, max_steps):
        n = self.count
        st = np.ascontiguousarray(stop_times, dtype=np.float64)
        ms = (C.c_longlong * n)(*[int(v) for v in max_steps])
        taken = (C.c_longlong * n)()
        cgt = (C.c_longlong * n)()
        cgp = (C.c_int * n)()
        last = (StepInfo * n)()
        status = (C.c_int * n)()
        rc = lib().pearl_batch_run(self._h, _ptr(st), ms, taken, cgt, cgp, last, status)
        out = dict(rc=rc, taken=list(taken), cg_total=list(cgt), cg_peak=list(cgp), last=list(last),
                   status=list(status))
        if rc != OK:
            out["error"] = lib().pearl_last_error(self._h).decode()
        return out
