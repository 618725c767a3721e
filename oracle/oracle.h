/* Oracle: a CPU restatement of Pearl's per-step solver loop (FP64).
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, the smoke check and the CPU-baseline
 * leg of bench.py load this library, and only as the checker.  The product
 * (the HIP library and its host) never links or calls it.
 *
 * Each function restates the reference function named in its comment
 * (file:line in pearl_serial.cpp), with the reference's operation order and
 * rounding (no FP contraction, fma exactly where the reference calls it).
 *
 * Reductions come in two orders:
 *   ORC_RED_SEQ      row-major sequential sums, as the serial reference
 *                    (array_dot pearl_serial.cpp:738-742, total_carbon :784-788,
 *                    reduced_energy :798-820).  Bitwise equal to pearl_serial.
 *   ORC_RED_BLOCKED  the canonical blocked order of the HIP kernels: tiles of
 *                    256 columns by k rows (red_k), numbered row of tiles by
 *                    row of tiles; thread t of a tile owns one column and adds
 *                    its k cells in increasing row; 64-lane halving tree per
 *                    wavefront; (w0+w1)+(w2+w3) per block; tile partials
 *                    reduced by one 256-thread block the same way (thread t
 *                    adds partials t, t+256, ...).
 * Max/min/AND reductions are order-independent and identical in both.
 *
 * Parity pins (see tests/): ORC_RED_SEQ results are compared bitwise with the
 * reference compiled from source (oracle/_ref), and with golden vectors in
 * tests/golden/ produced from it; closed-form and invariant checks cover the
 * pieces listed in the test files.
 *
 * Layout: every field is an unghosted row-major [ny][nx] plane, index
 * y*nx+x; phi is three such planes (alpha, theta, gamma).  x is periodic,
 * y has zero-flux faces.
 */
#ifndef PEARL_ORACLE_H
#define PEARL_ORACLE_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { ORC_RED_SEQ = 0, ORC_RED_BLOCKED = 1 };

/* Status codes; identical to the product C ABI (include/pearl_gpu.h). */
enum {
    ORC_OK = 0,
    ORC_ERR_INVALID_TIME = 1,   /* "invalid advance time"            :1121 */
    ORC_ERR_DT_MIN = 2,         /* "step rejected below dt_min: "+why :1171 */
    ORC_ERR_RETRY_LIMIT = 3,    /* "step retry limit: "+why          :1173 */
    ORC_ERR_MIXTURE = 4,        /* "Invalid mixture coefficients"     :136 */
    ORC_ERR_PROJECTION = 5,     /* "Simplex projection failed"        :505 */
    ORC_ERR_UNSAFE_SHIFT = 6,   /* "unsafe shift: removed strip contains gamma" :2068 */
    ORC_ERR_NONFINITE_CARBON = 7, /* "Nonfinite total carbon"         :143 */
    ORC_ERR_INVALID_SHIFT = 8,  /* "invalid shift"                    :2024 */
    ORC_ERR_ALLOC = 9
};

/* Rejection reasons ("why", pearl_serial.cpp:1132-1160). */
enum {
    ORC_WHY_PHASE = 0,        /* "phase increment" */
    ORC_WHY_LINEAR = 1,       /* "linear/closure" */
    ORC_WHY_BOUNDS = 2,       /* "composition/phase bounds" */
    ORC_WHY_MASS = 3,         /* "mass conservation" */
    ORC_WHY_ENERGY = 4        /* "free-energy increase" */
};

typedef struct {
    int nx, ny;
    double dx, dy;
    double A[3], B[3], C[3], D[3];           /* Thermo, :76-91 */
    double eps, sigma, triple;               /* InterfaceParameters, :431-442 */
    double tau[3];                           /* tau_at, tau_ag, tau_tg, :521-532 */
    double Vm;
    double dt_max, dt_min, max_phase_change; /* Controls, :767-773 */
    double cg_rtol, cg_atol, closure_tol;
    double energy_rtol, energy_atol;
    int cg_max, max_retries, check_energy, phase_only;
    int red_mode;                            /* ORC_RED_SEQ or ORC_RED_BLOCKED */
    int red_k;                               /* rows per tile = cells per thread (blocked order) */
} orc_params;

typedef struct {
    double *phi, *carbon, *mu;               /* 3n, n, n */
    double time, dt, shift_distance, initial_carbon, carbon_added, carbon_removed;
    long long step, shifts, rejected;
} orc_state;

typedef struct {
    double accepted_dt, max_phase_change, linear_residual, closure, energy_change;
    int cg_iterations, retries;
    int why;                                 /* reason of the last rejection */
} orc_stepinfo;

/* project_simplex :491-509 (0 on success). */
int orc_project_simplex(const double* trial, double* out);

/* Sum in the configured order. */
double orc_sum(const orc_params* P, const double* v, size_t n);

/* propose_phi :961-1002.  Returns max |phi'-phi| (inf on non-finite trial),
 * writes phi' (3n).  *status = ORC_ERR_PROJECTION if the projection fails. */
double orc_propose_phi(const orc_params* P, const double* phi, const double* mu, double dt,
                       double* phi_out, int* status);

/* implicit_diffusion :1010-1094.  Returns ORC_OK or ORC_ERR_MIXTURE; *ok,
 * *it, *residual, *closure as DiffusionInfo. */
int orc_implicit_diffusion(const orc_params* P, const double* phi_trial, const double* c_old,
                           const double* mu_old, double dt, double mu_ref, double* c_out,
                           double* mu_out, int* ok, int* it, double* residual, double* closure);

int orc_valid_state(const orc_params* P, const double* phi, const double* c, const double* mu); /* :1096 */
double orc_total_carbon(const orc_params* P, const double* c);                                 /* :784 */
double orc_reduced_energy(const orc_params* P, const double* phi, const double* c,
                          double mu_ref, int* status);                                          /* :793 */

/* advance :1119-1174.  Returns a status; on ORC_ERR_DT_MIN / RETRY_LIMIT the
 * reason is info->why. */
int orc_advance(const orc_params* P, orc_state* s, double mu_ref, double stop_time, orc_stepinfo* info);

/* close_mu :789-792 */
int orc_close_mu(const orc_params* P, orc_state* s);

/* maybe_shift :2053-2071 with shift_rows :2023-2052. */
int orc_maybe_shift(const orc_params* P, orc_state* s, int moving_window, double trigger,
                    double target, double c_gamma, int* shifted, double* front_out);

/* smooth_interface_only :869-958 (capillarity-only preprocessing).
 * report: steps, converged, energy_initial, energy_final, final_rel_change,
 * final_max_phase_change. */
int orc_smooth(const orc_params* P, double* phi, int max_steps, int min_steps, double energy_rtol,
               double max_phase_change, double* report);

#ifdef __cplusplus
}
#endif
#endif
