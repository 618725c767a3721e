/* pearl_gpu.h — C ABI of the MI355X (gfx950) per-step solver for Pearl.
 *
 * The reference has no plugin/FFI layer; its seam is the C++ call
 *   StepInfo advance(const Domain&, State&, const Thermo&, const PhaseParameters&,
 *                    const Controls&, double mu_reference, double stop_time)
 *   (pearl_serial.cpp:1119, called only from the step loop :2303)
 * plus maybe_shift (:2053, called at output times :2311).  This header
 * replaces that seam: one opaque context owns the device-resident State, and
 * each entry point below replaces the reference function named beside it.
 *
 * Conventions
 *  - Plain pointers and sizes; no C++ or torch types.  Host arrays are
 *    caller-owned and copied in/out (no aliasing).
 *  - Fields are row-major [ny][nx] planes (the layout of the HDF5 datasets,
 *    pearl_serial.cpp:1895-1929); phi is three planes alpha, theta, gamma.
 *  - Every call is synchronous with respect to the state; one context per
 *    host thread.  Return value: PEARL_OK (0) or a PEARL_ERR_* code; the
 *    message of the reference's exception is available from
 *    pearl_last_error().
 *  - The FP64 plain form computes every per-cell expression in the
 *    reference's order and rounding; global sums use a fixed blocked order
 *    (pearl_get_layout reports it), max/min/AND reductions are exact.
 */
#ifndef PEARL_GPU_H
#define PEARL_GPU_H

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes.  Messages are the reference's exception texts. */
#define PEARL_OK 0
#define PEARL_ERR_INVALID_TIME 1     /* "invalid advance time"                       :1121 */
#define PEARL_ERR_DT_MIN 2           /* "step rejected below dt_min: " + why          :1171 */
#define PEARL_ERR_RETRY_LIMIT 3      /* "step retry limit: " + why                    :1173 */
#define PEARL_ERR_MIXTURE 4          /* "Invalid mixture coefficients"                :136  */
#define PEARL_ERR_PROJECTION 5       /* "Simplex projection failed"                   :505  */
#define PEARL_ERR_UNSAFE_SHIFT 6     /* "unsafe shift: removed strip contains gamma"  :2068 */
#define PEARL_ERR_NONFINITE_CARBON 7 /* "Nonfinite total carbon"                      :143  */
#define PEARL_ERR_INVALID_SHIFT 8    /* "invalid shift"                               :2024 */
#define PEARL_ERR_ALLOC 9            /* device allocation failed */
#define PEARL_ERR_ARGUMENT 10        /* invalid argument (sizes, pointers, parameters) */
#define PEARL_ERR_DEVICE 11          /* HIP runtime error */
#define PEARL_ERR_SMOOTHING 12       /* "nonfinite smoothing force" :913 or
                                        "sharp smoothing could not find an energy-decreasing step" :941 */

/* Rejection reasons, the reference's "why" strings (:1132-1160). */
#define PEARL_WHY_PHASE 0   /* "phase increment" */
#define PEARL_WHY_LINEAR 1  /* "linear/closure" */
#define PEARL_WHY_BOUNDS 2  /* "composition/phase bounds" */
#define PEARL_WHY_MASS 3    /* "mass conservation" */
#define PEARL_WHY_ENERGY 4  /* "free-energy increase" */

/* Precision forms. */
#define PEARL_FP64 0        /* plain form: FP64 throughout */
#define PEARL_MIXED 1       /* mixed form: the Jacobi diagonal and the PCG vectors p and z stored
                               in FP32 and used exactly as stored; phi, c, mu, the operator's
                               coefficients (chi, K), x, r, Ap, every residual, dot product, the
                               carbon mass and the acceptance reductions in FP64 */
#define PEARL_FP32_STENCIL 2 /* the mixed form with FP32 stencil arithmetic in the phase proposal:
                               the force (obstacle, chemical, capillary face terms), its mean and
                               the step scale in FP32 (the chemical driving-force differences
                               formed in FP64 first); phi, the increment's application, the
                               projection and everything after it as in PEARL_MIXED */

/* Model and numerical parameters: the reference's Domain (:617), Thermo
 * (:76-91), PhaseParameters (:521-532) and Controls (:767-773). */
typedef struct {
    int nx, ny;                        /* global cells; x periodic, y zero flux; nx >= 3, ny >= 2,
                                          3*nx*ny < 2^31 - 2^20 */
    double dx, dy;                     /* m, each in [2^-75, 2^75] */
    double A[3], B[3], C[3], D[3];     /* f_p(c)=A c^2+B c+C [J/mol], 2A in [2^-150, 2^150];
                                          D_p [m^2/s] */
    double eps, sigma, triple;         /* interface parameters */
    double tau[3];                     /* tau_at, tau_ag, tau_tg [J s/m^4] */
    double Vm;                         /* m^3/mol, in [2^-150, 2^150] */
    double dt_max, dt_min, max_phase_change;
    double cg_rtol, cg_atol, closure_tol;
    double energy_rtol, energy_atol;
    int cg_max, max_retries, check_energy, phase_only;
    /* Extension (no reference counterpart): interface energy per phase pair
     * (alpha-theta, alpha-gamma, theta-gamma); 0 means `sigma`.  With every
     * entry 0 or equal to sigma the step is the reference's, bit for bit.
     * Otherwise the obstacle term is sum_{a<b} 16 sigma_ab/(pi^2 eps)
     * phi_a phi_b and the gradient term sum_{a<b} eps sigma_ab q_ab^2 (the
     * reference's forms with sigma per pair); the energy-test scale keeps
     * `sigma`.  A zero-initialised struct selects the reference. */
    double pair_sigma[3];
} pearl_params;

/* State scalars (State, :758-764). */
typedef struct {
    double time, dt, shift_distance, initial_carbon, carbon_added, carbon_removed;
    long long step, shifts, rejected;
} pearl_scalars;

/* StepInfo (:774-777) plus the reason of the last rejection. */
typedef struct {
    double accepted_dt, max_phase_change, linear_residual, closure, energy_change;
    int cg_iterations, retries, why;
} pearl_step_info;

/* DiffusionInfo (:1003-1007). */
typedef struct {
    int ok, it;
    double residual, closure;
} pearl_diffusion_info;

/* Output-time diagnostics of the reference's snapshot (:2266-2292). */
typedef struct {
    double c_min, c_max, mu_min, mu_max, phi_sum_error, mass, energy;
} pearl_diag;

/* One FILL operation of initialization = microsim (filling_phase_at
 * :1393-1514).  method: PEARL_FILL_*.  p: the operation's parameters in
 * input order, as read_config parsed and validated them
 * (parse_filling_operation :1209, validate_filling_operations :1251; unused entries
 * 0).  dir: cos and sin of each direction the operation uses, evaluated by
 * the caller with the reference's expression, cos(angle*pi/180.0) with the
 * literal pi = 3.14159265358979323846 (:1396): Y-junction arm k (angle
 * p[6+k]) in dir[k], the pearlite arm's and the ellipse's angle p[2] in
 * dir[0]; unused entries 0. */
#define PEARL_FILL_CUBE 0                /* FILLCUBE                  :1399 */
#define PEARL_FILL_CYLINDER 1            /* FILLCYLINDER              :1406 */
#define PEARL_FILL_CENTERBOX2D 2         /* FILLCENTERBOX2D           :1411 */
#define PEARL_FILL_YJUNCTION2D 3         /* FILLYJUNCTION2D           :1414 */
#define PEARL_FILL_YJUNCTIONLAMELLAE2D 4 /* FILLYJUNCTIONLAMELLAE2D   :1415 */
#define PEARL_FILL_ADHEREDPEARLITEARM2D 5 /* FILLADHEREDPEARLITEARM2D :1439 */
#define PEARL_FILL_PEARLITEELLIPSE2D 6   /* FILLPEARLITEELLIPSE2D     :1486 */
#define PEARL_FILL_PARAMS 16
typedef struct {
    int method, reserved;
    double p[PEARL_FILL_PARAMS];
    double dir[3][2];
} pearl_fill;

/* Composition modes of pearl_init_composition (composition_mode, :1725-1740). */
#define PEARL_COMPOSITION_MU 0   /* initial_mu or gamma_carbon: mu given */
#define PEARL_COMPOSITION_MEAN 1 /* mean_carbon: mu from the mean carbon */

/* Result of the capillarity-only smoothing (SmoothReport :843-852, without
 * the phase fractions). */
typedef struct {
    int steps, converged;
    double energy_initial, energy_final, final_relative_energy_change, final_max_phase_change;
} pearl_smooth_report;

typedef struct pearl_ctx pearl_ctx;

/* Create a context on HIP device `device` for one domain.  mu_reference is
 * initialize()'s return value (:1761, :2231).  cells_per_thread (1, 2, 4, 8
 * or 16; 0 = automatic) is the number of rows of a tile: each thread of a
 * cell kernel owns that many cells of one column.  It fixes the blocked
 * reduction order (pearl_get_layout).  Replaces the construction of
 * Domain/State (:2209, :2226). */
int pearl_create(const pearl_params* params, double mu_reference, int precision, int device,
                 int cells_per_thread, pearl_ctx** out);
int pearl_destroy(pearl_ctx* ctx);
const char* pearl_last_error(const pearl_ctx* ctx);

/* Upload / download the State (replaces load_global :725 and gather :695). */
int pearl_set_state(pearl_ctx* ctx, const double* phi, const double* carbon, const double* mu,
                    const pearl_scalars* scalars);
int pearl_get_state(pearl_ctx* ctx, double* phi, double* carbon, double* mu, pearl_scalars* scalars);

/* One accepted step (replaces advance, :1119-1174). */
int pearl_advance(pearl_ctx* ctx, double stop_time, pearl_step_info* info);

/* Up to max_steps accepted steps toward stop_time without host I/O between
 * them (the step loop :2296-2305 between two output times).  Stops early at
 * stop_time.  *taken receives the number of accepted steps; cg_total and
 * cg_peak accumulate as in the reference main loop.  The steps run under
 * device-side control: the host replays captured rounds of kernels and
 * synchronises only to see whether the run has finished. */
int pearl_run(pearl_ctx* ctx, double stop_time, long long max_steps, long long* taken,
              long long* cg_total, int* cg_peak, pearl_step_info* last);

/* Output-time tolerance of pearl_run and pearl_batch_run: a domain's run
 * ends after the first accepted step whose time reaches stop_time - eps.
 * The reference's loop takes a step to be at the output time when
 * time >= next_output - eps with eps = 1e-12*max(end_time, output_dt)
 * (:2307); with that eps, pearl_run(min(end_time, next_output), ...) ends
 * exactly where the reference's loop writes its next output, so the loop
 * between two outputs runs without a host round trip per step.  Default 0
 * (run until time >= stop_time).  Applies to every domain of the context. */
int pearl_set_stop_tolerance(pearl_ctx* ctx, double eps);

/* Moving window at an output time (replaces maybe_shift :2053-2071 and
 * shift_rows :2023-2052; close_mu :789).  front receives the detected front
 * position (-1 if none). */
int pearl_maybe_shift(pearl_ctx* ctx, int moving_window, double shift_trigger, double shift_target,
                      double shift_gamma_carbon, int* shifted, double* front);

/* Snapshot reductions (:2267-2282). */
int pearl_diagnostics(pearl_ctx* ctx, pearl_diag* out);

/* Snapshot reductions without any device allocation.  No counterpart in
 * the reference: for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  The same reductions and results as pearl_diagnostics, on a
 * state held in the caller's device memory: phi (three [ny][nx] planes,
 * alpha, theta, gamma), carbon and mu ([ny][nx]), on the current HIP device.
 * No context: the parameter block, the device scalars and the tile
 * partials live in one static device array per device (about 4.2 MB, held
 * until the process ends), so the domain may have at most 65,536 tiles of
 * 256 x cells_per_thread cells (0 = 4; 67 M cells at 4); larger domains are
 * rejected with PEARL_ERR_ARGUMENT.  The call allocates nothing, runs on
 * `stream` (a hipStream_t; NULL = the default stream) and synchronises it.
 * Calls that overlap on one device share that array and race.  Errors carry
 * their message in msg (msg_len bytes, may be NULL). */
int pearl_diagnostics_noalloc(const pearl_params* params, double mu_reference, int cells_per_thread,
                              const double* phi, const double* carbon, const double* mu, void* stream,
                              pearl_diag* out, char* msg, int msg_len);

/* Single reference functions on the current state, for parity tests.
 *   pearl_propose_phi        propose_phi :961 (phi_out: 3 planes, may be NULL)
 *   pearl_implicit_diffusion implicit_diffusion :1010 with a given trial phi
 *   pearl_reduced_energy     reduced_energy :793
 *   pearl_total_carbon       total_carbon :784
 *   pearl_valid_state        valid_state :1096 */
int pearl_propose_phi(pearl_ctx* ctx, double dt, double* phi_out, double* max_change);
int pearl_implicit_diffusion(pearl_ctx* ctx, const double* phi_trial, double dt, double* carbon_out,
                             double* mu_out, pearl_diffusion_info* info);
int pearl_reduced_energy(pearl_ctx* ctx, double* energy);
int pearl_total_carbon(pearl_ctx* ctx, double* mass);
int pearl_valid_state(pearl_ctx* ctx, int* valid);

/* Capillarity-only smoothing of the current phases in place (replaces
 * smooth_interface_only :869-958, which initialize runs on sharp FILL
 * labels, :1798): projected gradient steps on interface_energy_only
 * (:822-841) with up to 24 halvings of the step per iteration.  The
 * interface energy is summed in the blocked order.  Carbon and mu are left
 * as they are.  max_steps = 0 (with min_steps = 0) only evaluates the
 * interface energy of the current phases (energy_initial = energy_final,
 * converged = 1), as initialize reports it for the diffuse profiles
 * (:1802-1807).  Errors: PEARL_ERR_ARGUMENT ("invalid sharp-smoothing
 * controls"), PEARL_ERR_SMOOTHING, PEARL_ERR_PROJECTION. */
int pearl_smooth(pearl_ctx* ctx, int max_steps, int min_steps, double energy_rtol, double max_phase_change,
                 pearl_smooth_report* report);

/* Sharp phase labels of a FILL recipe on the device (replaces the labelling
 * loop of initialize, :1764-1795, for initialization = microsim, which calls
 * filling_phase_at :1393-1514 per cell; sharp_smooth's labels are two
 * FILLCUBE operations).  Every cell starts as gamma and the `count`
 * operations apply in order; the current phases become one-hot.  Carbon and
 * mu are left as they are.  Labels equal the reference's bit for bit; the one
 * per-cell transcendental, the sine of the pearlite arm's outline (:1474),
 * is the device's, so a cell whose |across| lies within an ulp of that
 * outline could differ.  Errors: PEARL_ERR_ARGUMENT ("invalid FILL
 * operation k": unknown method, a phase ID outside 0..2, a non-finite
 * parameter, or a pair count below 1). */
int pearl_fill_phases(pearl_ctx* ctx, const pearl_fill* ops, int count);

/* Kinds of pearl_init_profile. */
#define PEARL_PROFILE_LAMELLA 0 /* initialization = lamella :1767-1771 */
#define PEARL_PROFILE_FLAT 1    /* initialization = flat    :1787-1791 */
#define PEARL_PROFILE_CELLS 2   /* one value per cell: initialization = circle :1787-1791 */

/* The diffuse-profile initializations on the device (replaces the lamella,
 * flat and circle branches of initialize's loop, :1764-1795).  The profile
 * values come from the caller, because profile (:563-569) takes the C
 * library's sine; the device forms the three phases of every cell from them
 * with the reference's expressions:
 *   LAMELLA: col[nx] = profile of column x's signed distance to the theta
 *            band, row[ny] = profile(y - seed_height):
 *            phi = ((1 - col) row, col row, 1 - row)
 *   FLAT:    row[ny]: phi[phase_a] = row, phi[phase_b] = 1 - row, else 0
 *   CELLS:   cell[ny*nx] (row-major): phi[phase_a] = cell, phi[phase_b] =
 *            1 - cell, else 0
 * Arrays a kind does not use may be NULL.  Carbon and mu are left as they
 * are (pearl_init_composition follows).  Errors: PEARL_ERR_ARGUMENT
 * ("invalid profile": unknown kind, a missing array, a phase pair outside
 * 0..2 or equal, a non-finite value). */
int pearl_init_profile(pearl_ctx* ctx, int kind, const double* col, const double* row, const double* cell,
                       int phase_a, int phase_b);

/* Carbon and chemical potential of the current phases (replaces the end of
 * initialize, :1813-1836).  mode PEARL_COMPOSITION_MU: mu = value.  mode
 * PEARL_COMPOSITION_MEAN: value is the mean carbon and
 * mu = (nx*ny*value - sum intercept) / sum susceptibility (:1814-1822),
 * summed in the blocked order.  Then mu is uniform and c = composition(mu,
 * mixture(phi)) in every cell.  Returns mu (initialize's return value, the
 * run's mu_reference), the total carbon (total_carbon :784, blocked order;
 * State::initial_carbon) and valid_state (:1096-1116) of the result, which
 * initialize turns into "initial state out of model range".  The state
 * scalars are left as they are.  Errors: PEARL_ERR_ARGUMENT,
 * PEARL_ERR_MIXTURE ("Invalid mixture coefficients"). */
int pearl_init_composition(pearl_ctx* ctx, int mode, double value, double* mu, double* initial_carbon,
                           int* valid);

/* One step of size dt that is always taken: propose_phi (:961-1002), then
 * implicit_diffusion (:1010-1094) on the trial phases, then the trial state
 * becomes current -- advance() without the acceptance tests and the dt
 * controller.  The differentiable map of pearl_adjoint; with the same dt
 * sequence it reproduces an adaptive run's accepted steps.  Coupled mode,
 * FP64, domain 0.  No reference counterpart. */
int pearl_step_fixed(pearl_ctx* ctx, double dt);

/* `steps` such steps of sizes dts[0..steps-1] in one call: the device
 * controller runs them from replayed rounds (one synchronisation per batch
 * of rounds), with the same kernels and results as that many
 * pearl_step_fixed calls.  Single-domain FP64 coupled contexts.  No
 * reference counterpart. */
int pearl_run_fixed(pearl_ctx* ctx, int steps, const double* dts);

/* Gradient of the loss
 *   L = 1/2 sum_cells,phases (phi_N - phi_target)^2 + 1/2 w_front (H - H_target)^2,
 *   H = dy/nx * sum_cells (1 - phi_gamma)   (the transformed height),
 * after `steps` fixed steps of sizes dts[0..steps-1] from the current state,
 * with respect to theta = {sigma, tau_at, tau_ag, tau_tg, D_alpha, D_theta,
 * D_gamma, sigma_at, sigma_ag, sigma_tg} (grad[10]; grad[0] moves the three
 * pair energies together and is the sum of grad[7..9]).  A hand-written discrete adjoint of the step: the
 * forward keeps a checkpoint every `every` steps; the reverse sweep
 * recomputes each segment from its checkpoint and runs the adjoint of the
 * diffusion solve (one PCG solve with the same operator) and of the
 * projected phase update.  The diffusion's converged solution is
 * differentiated as an exact solve (mu' = mu_ref + x), so the gradient is
 * exact to the PCG tolerance.  On return the context holds the final state.
 * Coupled mode, FP64, single-domain contexts.  No reference counterpart
 * (the reference has no derivatives). */
int pearl_adjoint(pearl_ctx* ctx, int steps, const double* dts, int every, const double* phi_target,
                  double w_front, double front_target, double* loss, double* grad);

/* Blocked-reduction layout in use: threads per block (256) and cells per
 * thread k.  A sum over an nx-by-ny domain is: tiles of 256 columns by k
 * rows, numbered row of tiles by row of tiles (b = ty*ceil(nx/256) + tx);
 * thread t of tile b owns column 256*tx + t and adds its cells of the tile
 * in increasing row index; a 64-lane halving tree per wavefront;
 * (w0+w1)+(w2+w3) per block; the tile partials are reduced by one
 * 256-thread block in the same way (thread t adds partials t, t+256, ...). */
int pearl_get_layout(const pearl_ctx* ctx, int* threads_per_block, int* cells_per_thread);

/* Per-kernel timing with HIP events on the context's stream.  Kernel ids:
 * see PEARL_KERNEL_*.  When enabled, every launch is bracketed by events;
 * pearl_kernel_stats returns the launch count, total milliseconds and the
 * algorithmic HBM bytes of all launches of that kernel.  Launches that a
 * device-side decision turned into no-ops (a rejected phase increment, a
 * PCG solve that had already finished) are counted under
 * PEARL_KERNEL_CG_IDLE with zero bytes. */
#define PEARL_KERNEL_PROPOSE 0        /* propose_phi */
#define PEARL_KERNEL_PREPARE 1        /* diffusion setup + initial residual, diagonal, tolerance */
#define PEARL_KERNEL_CG_APPLY 2       /* direction update + A p + p.Ap + the previous step x += alpha p */
#define PEARL_KERNEL_CG_UPDATE 3      /* r, z updates + r.z, r.r */
#define PEARL_KERNEL_TRUE_RESIDUAL 4  /* true residual after a failed check inside finalize (+ restart) */
#define PEARL_KERNEL_CG_IDLE 5        /* speculative PCG launches that found the solve finished */
#define PEARL_KERNEL_FINALIZE 6       /* flux update, closure, mu + acceptance reductions */
#define PEARL_KERNEL_STATE_REDUCE 7   /* valid/mass/energy of the current state, diagnostics */
#define PEARL_KERNEL_WINDOW 8         /* moving-window front, strip, shift */
#define PEARL_KERNEL_FINISH 9         /* one-block finishers of the blocked reductions */
#define PEARL_KERNEL_CG_FIRST 10      /* first direction of a solve, p = z: A p + p.Ap (no x step) */
#define PEARL_KERNEL_COUNT 11
int pearl_profile(pearl_ctx* ctx, int enable);
int pearl_kernel_stats(pearl_ctx* ctx, int kernel, long long* launches, double* total_ms, double* bytes);
const char* pearl_kernel_name(int kernel);

/* Batched ensemble: `count` independent domains of any sizes in one
 * context (one packed HBM arena; one launch of each kernel serves all of
 * them).  There is no reference counterpart; each domain's step is the
 * reference's advance.  pearl_batch_run advances domain d until it has taken
 * max_steps[d] accepted steps or reached stop_times[d]; the arrays taken,
 * cg_total (accumulated), cg_peak, last and status have `count` entries and
 * may be NULL.  Returns the first failing domain's status (message in
 * pearl_last_error); the others keep running to their own targets.  The
 * single-domain calls above act on domain 0.  The domains' tiles together
 * (256 x tile_rows cells each) are at most 2^32 / 512, one launch grid. */
int pearl_batch_create(int count, const pearl_params* params, const double* mu_references, int precision,
                       int device, int cells_per_thread, pearl_ctx** out);
int pearl_batch_size(const pearl_ctx* ctx, int* count);
int pearl_batch_set_state(pearl_ctx* ctx, int domain, const double* phi, const double* carbon, const double* mu,
                          const pearl_scalars* scalars);
int pearl_batch_get_state(pearl_ctx* ctx, int domain, double* phi, double* carbon, double* mu,
                          pearl_scalars* scalars);
int pearl_batch_run(pearl_ctx* ctx, const double* stop_times, const long long* max_steps, long long* taken,
                    long long* cg_total, int* cg_peak, pearl_step_info* last, int* status);

/* Attempt/iteration counters since creation (summed over domains):
 * attempts, rejected attempts, total CG iterations over all attempts. */
int pearl_counters(const pearl_ctx* ctx, long long* attempts, long long* rejected_attempts,
                   long long* cg_iterations);

/* Measurement utility, no reference counterpart: the bandwidth of a plain
 * device copy of `bytes` bytes (a multiple of 16, at most 32 GiB) on HIP
 * device `device`, by a hand-written kernel that moves one 16-byte vector
 * per thread, with and without the nontemporal hint (the faster counts).
 * `reps` timed copies after three warm-up copies, HIP events on a stream of
 * its own: *gbps = 2 * bytes * reps / seconds / 1e9 (each copy reads and
 * writes `bytes`).  The measured ceiling reported beside the vendor peak. */
int pearl_stream_probe(int device, long long bytes, int reps, double* gbps);

/* Measurement utility, no reference counterpart: the rate a cell kernel's
 * access pattern can reach without its arithmetic.  A kernel reads `reads`
 * and writes `writes` FP64 planes of nx*ny cells (1..8 each), nontemporal,
 * one column of rows_per_thread (1, 2, 4, 8, 16) cells per thread over
 * 256-column tiles, in one allocation with the planes spaced as in a
 * context's arena, as the cell kernels do.  `reps` timed launches after
 * three warm-up launches, HIP events on a stream of its own: *us per launch
 * and *gbps = (reads + writes) * 8 * nx * ny / seconds / 1e9.  Allocates
 * the planes for the call. */
int pearl_stream_mix(int device, int reads, int writes, int nx, int ny, int rows_per_thread, int reps, double* us,
                     double* gbps);

#ifdef __cplusplus
}
#endif
#endif
