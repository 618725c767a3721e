"""ctypes bridge to the oracle restatement (liboracle.so) and, where it was
built, to the reference harness (_ref/libpearl_ref.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, the smoke check and the
CPU-baseline leg of bench.py, as the checker.  The product never imports it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_PATH = os.path.join(HERE, "_ref", "libpearl_ref.so")

RED_SEQ, RED_BLOCKED = 0, 1

WHY = ["phase increment", "linear/closure", "composition/phase bounds",
       "mass conservation", "free-energy increase"]

_dp = C.POINTER(C.c_double)


class Params(C.Structure):
    _fields_ = [
        ("nx", C.c_int), ("ny", C.c_int), ("dx", C.c_double), ("dy", C.c_double),
        ("A", C.c_double * 3), ("B", C.c_double * 3), ("C", C.c_double * 3), ("D", C.c_double * 3),
        ("eps", C.c_double), ("sigma", C.c_double), ("triple", C.c_double),
        ("tau", C.c_double * 3), ("Vm", C.c_double),
        ("dt_max", C.c_double), ("dt_min", C.c_double), ("max_phase_change", C.c_double),
        ("cg_rtol", C.c_double), ("cg_atol", C.c_double), ("closure_tol", C.c_double),
        ("energy_rtol", C.c_double), ("energy_atol", C.c_double),
        ("cg_max", C.c_int), ("max_retries", C.c_int), ("check_energy", C.c_int),
        ("phase_only", C.c_int), ("red_mode", C.c_int), ("red_k", C.c_int),
    ]


class State(C.Structure):
    _fields_ = [
        ("phi", _dp), ("carbon", _dp), ("mu", _dp),
        ("time", C.c_double), ("dt", C.c_double), ("shift_distance", C.c_double),
        ("initial_carbon", C.c_double), ("carbon_added", C.c_double), ("carbon_removed", C.c_double),
        ("step", C.c_longlong), ("shifts", C.c_longlong), ("rejected", C.c_longlong),
    ]


class StepInfo(C.Structure):
    _fields_ = [
        ("accepted_dt", C.c_double), ("max_phase_change", C.c_double),
        ("linear_residual", C.c_double), ("closure", C.c_double), ("energy_change", C.c_double),
        ("cg_iterations", C.c_int), ("retries", C.c_int), ("why", C.c_int),
    ]


def build():
    subprocess.check_call(["make", "-s", "-C", HERE, "liboracle.so"])


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        P = C.POINTER(Params)
        L.orc_project_simplex.restype = C.c_int
        L.orc_project_simplex.argtypes = [_dp, _dp]
        L.orc_sum.restype = C.c_double
        L.orc_sum.argtypes = [P, _dp, C.c_size_t]
        L.orc_propose_phi.restype = C.c_double
        L.orc_propose_phi.argtypes = [P, _dp, _dp, C.c_double, _dp, C.POINTER(C.c_int)]
        L.orc_implicit_diffusion.restype = C.c_int
        L.orc_implicit_diffusion.argtypes = [P, _dp, _dp, _dp, C.c_double, C.c_double, _dp, _dp,
                                             C.POINTER(C.c_int), C.POINTER(C.c_int), _dp, _dp]
        L.orc_valid_state.restype = C.c_int
        L.orc_valid_state.argtypes = [P, _dp, _dp, _dp]
        L.orc_total_carbon.restype = C.c_double
        L.orc_total_carbon.argtypes = [P, _dp]
        L.orc_reduced_energy.restype = C.c_double
        L.orc_reduced_energy.argtypes = [P, _dp, _dp, C.c_double, C.POINTER(C.c_int)]
        L.orc_advance.restype = C.c_int
        L.orc_advance.argtypes = [P, C.POINTER(State), C.c_double, C.c_double, C.POINTER(StepInfo)]
        L.orc_close_mu.restype = C.c_int
        L.orc_close_mu.argtypes = [P, C.POINTER(State)]
        L.orc_maybe_shift.restype = C.c_int
        L.orc_maybe_shift.argtypes = [P, C.POINTER(State), C.c_int, C.c_double, C.c_double, C.c_double,
                                      C.POINTER(C.c_int), _dp]
        L.orc_smooth.restype = C.c_int
        L.orc_smooth.argtypes = [P, _dp, C.c_int, C.c_int, C.c_double, C.c_double, _dp]
        _lib = L
    return _lib


def _ptr(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_dp)


# Parameter vector layout shared with the reference harness (ref_params).
PARAM_KEYS = ["A0", "B0", "C0", "D0", "A1", "B1", "C1", "D1", "A2", "B2", "C2", "D2",
              "eps", "sigma", "triple", "tau_at", "tau_ag", "tau_tg", "Vm", "dx", "dy", "nx", "ny",
              "dt_max", "dt_min", "max_phase_change", "cg_rtol", "cg_atol", "closure_tol",
              "energy_rtol", "energy_atol", "cg_max", "max_retries", "check_energy", "phase_only",
              "dt", "end_time", "output_dt", "moving_window", "shift_trigger", "shift_target",
              "shift_gamma_carbon", "max_steps", "initial_mu", "T", "initial_carbon"]


def params_from_vector(v, red_mode=RED_SEQ, red_k=1) -> Params:
    d = dict(zip(PARAM_KEYS, [float(x) for x in v]))
    return params_from_dict(d, red_mode, red_k)


def params_from_dict(d: dict, red_mode=RED_SEQ, red_k=1) -> Params:
    p = Params()
    p.nx, p.ny = int(d["nx"]), int(d["ny"])
    p.dx, p.dy = d["dx"], d["dy"]
    for i in range(3):
        p.A[i], p.B[i], p.C[i], p.D[i] = d[f"A{i}"], d[f"B{i}"], d[f"C{i}"], d[f"D{i}"]
    p.eps, p.sigma, p.triple = d["eps"], d["sigma"], d["triple"]
    p.tau[0], p.tau[1], p.tau[2] = d["tau_at"], d["tau_ag"], d["tau_tg"]
    p.Vm = d["Vm"]
    for k in ("dt_max", "dt_min", "max_phase_change", "cg_rtol", "cg_atol", "closure_tol",
              "energy_rtol", "energy_atol"):
        setattr(p, k, d[k])
    p.cg_max, p.max_retries = int(d["cg_max"]), int(d["max_retries"])
    p.check_energy, p.phase_only = int(d["check_energy"]), int(d["phase_only"])
    p.red_mode, p.red_k = red_mode, red_k
    return p


class OracleState:
    """Owns numpy arrays for a State (planar phi [3,ny,nx], carbon, mu)."""

    def __init__(self, phi, carbon, mu, scalars=None):
        self.phi = np.ascontiguousarray(phi, dtype=np.float64).copy()
        self.carbon = np.ascontiguousarray(carbon, dtype=np.float64).copy()
        self.mu = np.ascontiguousarray(mu, dtype=np.float64).copy()
        self.st = State()
        self.st.phi, self.st.carbon, self.st.mu = _ptr(self.phi), _ptr(self.carbon), _ptr(self.mu)
        if scalars is not None:
            self.set_scalars(scalars)

    SCALARS = ["time", "dt", "shift_distance", "initial_carbon", "carbon_added", "carbon_removed",
               "step", "shifts", "rejected"]

    def set_scalars(self, sc):
        for k, v in zip(self.SCALARS, sc):
            setattr(self.st, k, int(v) if k in ("step", "shifts", "rejected") else float(v))

    def scalars(self):
        return [float(getattr(self.st, k)) for k in self.SCALARS]


def propose_phi(P: Params, phi, mu, dt):
    out = np.empty_like(phi)
    st = C.c_int(0)
    ch = lib().orc_propose_phi(C.byref(P), _ptr(phi), _ptr(mu), dt, _ptr(out), C.byref(st))
    return ch, out, st.value


def implicit_diffusion(P: Params, phi_t, c_old, mu_old, dt, mu_ref):
    c = np.empty_like(c_old)
    m = np.empty_like(mu_old)
    ok, it = C.c_int(0), C.c_int(0)
    res, clo = C.c_double(0), C.c_double(0)
    st = lib().orc_implicit_diffusion(C.byref(P), _ptr(phi_t), _ptr(c_old), _ptr(mu_old), dt, mu_ref,
                                      _ptr(c), _ptr(m), C.byref(ok), C.byref(it), C.byref(res), C.byref(clo))
    return dict(status=st, ok=ok.value, it=it.value, residual=res.value, closure=clo.value, carbon=c, mu=m)


def valid_state(P, phi, c, mu):
    return lib().orc_valid_state(C.byref(P), _ptr(phi), _ptr(c), _ptr(mu))


def total_carbon(P, c):
    return lib().orc_total_carbon(C.byref(P), _ptr(c))


def reduced_energy(P, phi, c, mu_ref):
    st = C.c_int(0)
    e = lib().orc_reduced_energy(C.byref(P), _ptr(phi), _ptr(c), mu_ref, C.byref(st))
    return e, st.value


def project_simplex(z):
    z = np.ascontiguousarray(z, dtype=np.float64)
    out = np.zeros(3)
    rc = lib().orc_project_simplex(_ptr(z), _ptr(out))
    return rc, out


def osum(P, v):
    v = np.ascontiguousarray(v, dtype=np.float64)
    return lib().orc_sum(C.byref(P), _ptr(v), v.size)


def advance(P: Params, s: OracleState, mu_ref, stop_time):
    info = StepInfo()
    rc = lib().orc_advance(C.byref(P), C.byref(s.st), mu_ref, stop_time, C.byref(info))
    return rc, info


def maybe_shift(P: Params, s: OracleState, on, trigger, target, c_gamma):
    sh = C.c_int(0)
    fr = C.c_double(0)
    rc = lib().orc_maybe_shift(C.byref(P), C.byref(s.st), on, trigger, target, c_gamma, C.byref(sh), C.byref(fr))
    return rc, sh.value, fr.value


def smooth(P: Params, phi, max_steps, min_steps, rtol, max_change):
    rep = np.zeros(6)
    rc = lib().orc_smooth(C.byref(P), _ptr(phi), max_steps, min_steps, rtol, max_change, _ptr(rep))
    return rc, rep


def run(P: Params, s: OracleState, mu_ref, end_time, output_dt, max_steps, window=None):
    """The reference main loop (pearl_serial.cpp:2254, :2296-2316) without
    file output.  window = (trigger, target, c_gamma) or None."""
    import math
    next_output = (math.floor((s.st.time + 1e-12 * output_dt) / output_dt) + 1.0) * output_dt
    taken = cg_total = cg_peak = shifts = 0
    while s.st.time < end_time and taken < max_steps:
        rc, info = advance(P, s, mu_ref, min(end_time, next_output))
        if rc != 0:
            return rc, dict(taken=taken, cg_total=cg_total, cg_peak=cg_peak, shifts=shifts, why=info.why)
        taken += 1
        cg_total += info.cg_iterations
        cg_peak = max(cg_peak, info.cg_iterations)
        eps = 1e-12 * max(end_time, output_dt)
        if s.st.time >= next_output - eps or s.st.time >= end_time - eps:
            if abs(s.st.time - next_output) <= eps:
                s.st.time = next_output
            if abs(s.st.time - end_time) <= eps:
                s.st.time = end_time
            if window is not None:
                rc, sh, _ = maybe_shift(P, s, 1, *window)
                if rc != 0:
                    return rc, dict(taken=taken, cg_total=cg_total, cg_peak=cg_peak, shifts=shifts, why=-1)
                shifts += sh
            next_output += output_dt
    return 0, dict(taken=taken, cg_total=cg_total, cg_peak=cg_peak, shifts=shifts, why=-1)


# ---------------------------------------------------------------------------
# Reference harness (the reference compiled from its own source)
# ---------------------------------------------------------------------------
_ref = None


def ref_available() -> bool:
    return os.path.exists(REF_PATH)


def ref_lib():
    global _ref
    if _ref is None:
        L = C.CDLL(REF_PATH)
        vp = C.c_void_p
        L.ref_open.restype = vp
        L.ref_open.argtypes = [C.c_char_p, C.c_char_p, C.c_int]
        L.ref_close.argtypes = [vp]
        L.ref_last_error.restype = C.c_char_p
        L.ref_last_error.argtypes = [vp]
        L.ref_params.argtypes = [vp, _dp]
        L.ref_signature.argtypes = [vp, C.c_char_p, C.c_int]
        L.ref_initialize.argtypes = [vp]
        L.ref_sharp_labels.argtypes = [vp, _dp]
        L.ref_get_state.argtypes = [vp, _dp, _dp, _dp, _dp]
        L.ref_set_state.argtypes = [vp, _dp, _dp, _dp, _dp]
        L.ref_set_controls.argtypes = [vp, _dp]
        L.ref_advance.argtypes = [vp, C.c_double, _dp]
        L.ref_run.argtypes = [vp, C.c_double, C.c_longlong, _dp]
        L.ref_propose_phi.restype = C.c_double
        L.ref_propose_phi.argtypes = [vp, C.c_double, _dp]
        L.ref_implicit_diffusion.argtypes = [vp, _dp, C.c_double, _dp, _dp, _dp]
        L.ref_reduced_energy.restype = C.c_double
        L.ref_reduced_energy.argtypes = [vp]
        L.ref_total_carbon.restype = C.c_double
        L.ref_total_carbon.argtypes = [vp]
        L.ref_valid_state.argtypes = [vp]
        L.ref_maybe_shift.argtypes = [vp, C.POINTER(C.c_int)]
        L.ref_set_window.argtypes = [vp, C.c_int, C.c_double, C.c_double, C.c_double]
        L.ref_smooth.argtypes = [vp, C.c_int, C.c_int, C.c_double, C.c_double, _dp]
        L.ref_write_frame.argtypes = [vp, C.c_char_p]
        L.ref_read_checkpoint.argtypes = [vp, C.c_char_p]
        _ref = L
    return _ref


class Ref:
    """The reference solver on one parsed input file."""

    def __init__(self, path: str):
        L = ref_lib()
        err = C.create_string_buffer(512)
        self.h = L.ref_open(path.encode(), err, 512)
        if not self.h:
            raise RuntimeError(err.value.decode())
        v = np.zeros(64)
        L.ref_params(self.h, _ptr(v))
        self.pvec = v[:len(PARAM_KEYS)].copy()
        self.pd = dict(zip(PARAM_KEYS, self.pvec.tolist()))
        self.nx, self.ny = int(self.pd["nx"]), int(self.pd["ny"])

    def __del__(self):
        if getattr(self, "h", None):
            ref_lib().ref_close(self.h)
            self.h = None

    def _chk(self, rc):
        if rc != 0:
            raise RuntimeError(ref_lib().ref_last_error(self.h).decode())

    def signature(self) -> str:
        buf = C.create_string_buffer(1 << 16)
        n = ref_lib().ref_signature(self.h, buf, 1 << 16)
        assert n >= 0
        return buf.value.decode()

    def initialize(self):
        self._chk(ref_lib().ref_initialize(self.h))

    def sharp_labels(self):
        """phi[3, ny, nx] of the reference's initialize at its init_sharp
        observer call (before smoothing); the handle must be reopened
        afterwards."""
        phi = np.empty((3, self.ny, self.nx))
        self._chk(ref_lib().ref_sharp_labels(self.h, _ptr(phi)))
        return phi

    def get_state(self):
        n = self.nx * self.ny
        phi = np.zeros((3, self.ny, self.nx))
        c = np.zeros((self.ny, self.nx))
        mu = np.zeros((self.ny, self.nx))
        sc = np.zeros(10)
        ref_lib().ref_get_state(self.h, _ptr(phi), _ptr(c), _ptr(mu), _ptr(sc))
        assert phi.size == 3 * n
        return phi, c, mu, sc

    def set_state(self, phi, c, mu, sc):
        phi = np.ascontiguousarray(phi, dtype=np.float64)
        c = np.ascontiguousarray(c, dtype=np.float64)
        mu = np.ascontiguousarray(mu, dtype=np.float64)
        sc = np.ascontiguousarray(sc, dtype=np.float64)
        ref_lib().ref_set_state(self.h, _ptr(phi), _ptr(c), _ptr(mu), _ptr(sc))

    def set_controls(self, **kw):
        keys = ["dt_max", "dt_min", "max_phase_change", "cg_rtol", "cg_atol", "closure_tol",
                "energy_rtol", "energy_atol", "cg_max", "max_retries"]
        cur = {k: self.pd[k] for k in keys}
        cur.update(kw)
        self.pd.update(kw)
        v = np.array([cur[k] for k in keys], dtype=np.float64)
        ref_lib().ref_set_controls(self.h, _ptr(v))

    def advance(self, stop_time):
        info = np.zeros(8)
        rc = ref_lib().ref_advance(self.h, stop_time, _ptr(info))
        if rc != 0:
            return None, ref_lib().ref_last_error(self.h).decode()
        return info, None

    def run(self, end_time, max_steps):
        st = np.zeros(4)
        self._chk(ref_lib().ref_run(self.h, end_time, max_steps, _ptr(st)))
        return dict(taken=int(st[0]), cg_total=int(st[1]), cg_peak=int(st[2]), shifts=int(st[3]))

    def propose_phi(self, dt):
        out = np.zeros((3, self.ny, self.nx))
        ch = ref_lib().ref_propose_phi(self.h, dt, _ptr(out))
        return ch, out

    def implicit_diffusion(self, phi_t, dt):
        phi_t = np.ascontiguousarray(phi_t, dtype=np.float64)
        c = np.zeros((self.ny, self.nx))
        m = np.zeros((self.ny, self.nx))
        info = np.zeros(4)
        self._chk(ref_lib().ref_implicit_diffusion(self.h, _ptr(phi_t), dt, _ptr(c), _ptr(m), _ptr(info)))
        return dict(ok=int(info[0]), it=int(info[1]), residual=info[2], closure=info[3], carbon=c, mu=m)

    def reduced_energy(self):
        return ref_lib().ref_reduced_energy(self.h)

    def total_carbon(self):
        return ref_lib().ref_total_carbon(self.h)

    def valid_state(self):
        return ref_lib().ref_valid_state(self.h)

    def set_window(self, on, trigger, target, c_gamma):
        ref_lib().ref_set_window(self.h, on, trigger, target, c_gamma)

    def maybe_shift(self):
        sh = C.c_int(0)
        rc = ref_lib().ref_maybe_shift(self.h, C.byref(sh))
        if rc != 0:
            return None, ref_lib().ref_last_error(self.h).decode()
        return sh.value, None

    def smooth(self, max_steps, min_steps, rtol, max_change):
        rep = np.zeros(6)
        self._chk(ref_lib().ref_smooth(self.h, max_steps, min_steps, rtol, max_change, _ptr(rep)))
        return rep

    def write_frame(self, path):
        self._chk(ref_lib().ref_write_frame(self.h, path.encode()))

    def read_checkpoint(self, path):
        self._chk(ref_lib().ref_read_checkpoint(self.h, path.encode()))
