/* Oracle restatement of Pearl's per-step solver loop.  TEST INFRASTRUCTURE:
 * the checker for the HIP path, never part of it.  See oracle.h for the
 * layout, the two reduction orders and what pins each function.
 *
 * Every expression keeps the reference's evaluation order (pearl_serial.cpp
 * lines cited per function); compile with -ffp-contract=off.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define RED_THREADS 256

/* ------------------------------------------------------------------ */
/* Reductions                                                          */
/* ------------------------------------------------------------------ */

/* One 256-lane block: per-wavefront halving tree (lane l += lane l+s for
 * s = 32..1; equal to an xor butterfly because IEEE addition commutes),
 * then (w0+w1)+(w2+w3). */
static double block_tree(double* a) {
    double w[4];
    for (int wv = 0; wv < 4; ++wv) {
        double* v = a + 64 * wv;
        for (int s = 32; s >= 1; s >>= 1)
            for (int l = 0; l < s; ++l) v[l] = v[l] + v[l + s];
        w[wv] = v[0];
    }
    return (w[0] + w[1]) + (w[2] + w[3]);
}

/* Sum of a row-major [rows][nx] array in the canonical blocked order: tiles
 * of 256 columns by ty rows, numbered row of tiles by row of tiles
 * (b = tile_row*tiles_x + tile_col); thread t of tile b owns column
 * 256*tile_col + t and adds its cells of the tile in increasing row; then the
 * block tree; then the tile partials reduced by one 256-thread block the
 * same way (thread t adds partials t, t+256, ...). */
static double blocked_sum(const double* v, size_t n, int nx, int ty) {
    if (ty < 1) ty = 1;
    if (nx < 1 || n % (size_t)nx != 0) nx = (int)(n == 0 ? 1 : n);
    const size_t rows = n / (size_t)nx;
    const size_t tiles_x = ((size_t)nx + RED_THREADS - 1) / RED_THREADS;
    const size_t tiles_y = rows == 0 ? 1 : (rows + (size_t)ty - 1) / (size_t)ty;
    const size_t nb = tiles_x * tiles_y;
    double* part = (double*)malloc(nb * sizeof(double));
    double acc[RED_THREADS];
    for (size_t b = 0; b < nb; ++b) {
        const size_t tx = b % tiles_x, y0 = (b / tiles_x) * (size_t)ty;
        for (int t = 0; t < RED_THREADS; ++t) {
            const size_t x = tx * RED_THREADS + (size_t)t;
            double s = 0.0;
            if (x < (size_t)nx)
                for (size_t y = y0; y < y0 + (size_t)ty && y < rows; ++y) s += v[y * (size_t)nx + x];
            acc[t] = s;
        }
        part[b] = block_tree(acc);
    }
    for (int t = 0; t < RED_THREADS; ++t) {
        double s = 0.0;
        for (size_t j = (size_t)t; j < nb; j += RED_THREADS) s += part[j];
        acc[t] = s;
    }
    free(part);
    return block_tree(acc);
}

double orc_sum(const orc_params* P, const double* v, size_t n) {
    if (P->red_mode == ORC_RED_BLOCKED) return blocked_sum(v, n, P->nx, P->red_k);
    double s = 0;
    for (size_t i = 0; i < n; ++i) s += v[i];
    return s;
}

/* array_dot :738-742 */
static double dot(const orc_params* P, const double* a, const double* b, size_t n, double* tmp) {
    if (P->red_mode == ORC_RED_BLOCKED) {
        for (size_t i = 0; i < n; ++i) tmp[i] = a[i] * b[i];
        return blocked_sum(tmp, n, P->nx, P->red_k);
    }
    double s = 0;
    for (size_t i = 0; i < n; ++i) s += a[i] * b[i];
    return s;
}

static double dmax(double a, double b) { return (a < b) ? b : a; } /* std::max */
static double dmin(double a, double b) { return (b < a) ? b : a; } /* std::min */

/* ------------------------------------------------------------------ */
/* Pointwise thermodynamics :76-145                                    */
/* ------------------------------------------------------------------ */

static double psi(const orc_params* P, int p, double mu) { /* Parabola::psi :85-89 */
    const double d = mu - P->B[p];
    return fma(-0.25 / P->A[p], d * d, P->C[p]);
}

static void weights(const double* phi, double* h) { /* :102-107 */
    const double triple = 2.0 * phi[0] * phi[1] * phi[2];
    for (int p = 0; p < 3; ++p) h[p] = phi[p] * phi[p] * (3.0 - 2.0 * phi[p]) + triple;
}

typedef struct { double s, c0, K; } mix_t;

static int mixture(const orc_params* P, const double* phi, mix_t* m) { /* :125-138 */
    double h[3];
    weights(phi, h);
    m->s = 0.0;
    m->c0 = 0.0;
    m->K = 0.0;
    for (int p = 0; p < 3; ++p) {
        const double chi = 1.0 / (2.0 * P->A[p]);
        m->s += h[p] * chi;
        m->c0 -= h[p] * chi * P->B[p];
        m->K += phi[p] * P->D[p] * chi;
    }
    if (!(m->s > 0.0) || !isfinite(m->s) || !isfinite(m->c0) || !isfinite(m->K) || m->K < 0.0)
        return ORC_ERR_MIXTURE;
    return ORC_OK;
}

static double composition(double mu, const mix_t* m) { return fma(m->s, mu, m->c0); } /* :139-141 */

static double local_tau(const double* p, const double* tau) { /* :533-537 */
    const double w0 = p[0] * p[1], w1 = p[0] * p[2], w2 = p[1] * p[2];
    const double s = w0 + w1 + w2;
    if (s > 1e-30) return (w0 * tau[0] + w1 * tau[1] + w2 * tau[2]) / s;
    double m = tau[0]; /* std::min_element: first smallest */
    if (tau[1] < m) m = tau[1];
    if (tau[2] < m) m = tau[2];
    return m;
}

/* project_simplex :491-509.  Returns 0 on success. */
static int project_simplex(const double* trial, double* out) {
    double s[3] = {trial[0], trial[1], trial[2]};
    /* descending sort of three values */
    if (s[0] < s[1]) { double t = s[0]; s[0] = s[1]; s[1] = t; }
    if (s[1] < s[2]) { double t = s[1]; s[1] = s[2]; s[2] = t; }
    if (s[0] < s[1]) { double t = s[0]; s[0] = s[1]; s[1] = t; }
    double sum = 0.0, threshold = 0.0;
    int active = 0;
    for (int j = 0; j < 3; ++j) {
        sum += s[j];
        const double candidate = (sum - 1.0) / (j + 1);
        if (s[j] > candidate) {
            active = j + 1;
            threshold = candidate;
        }
    }
    if (active == 0) return ORC_ERR_PROJECTION;
    for (int a = 0; a < 3; ++a) out[a] = dmax(0.0, trial[a] - threshold);
    return ORC_OK;
}

/* Exported for the brute-force pin of the projection. */
int orc_project_simplex(const double* trial, double* out) { return project_simplex(trial, out); }

static inline void cell_phi(const double* phi, size_t n, size_t i, double* p) {
    p[0] = phi[i];
    p[1] = phi[n + i];
    p[2] = phi[2 * n + i];
}

static double pi_value(void) { return acos(-1.0); }

/* ------------------------------------------------------------------ */
/* propose_phi :961-1002                                               */
/* ------------------------------------------------------------------ */
double orc_propose_phi(const orc_params* P, const double* phi, const double* mu, double dt,
                       double* phi_out, int* status) {
    const int nx = P->nx, ny = P->ny;
    const size_t n = (size_t)nx * ny;
    const double eps = P->eps, sigma = P->sigma;
    const double obs = 16 * sigma / (pi_value() * pi_value() * eps), tr = P->triple / eps;
    double change = 0;
    *status = ORC_OK;
    for (int y = 0; y < ny; ++y)
        for (int x = 0; x < nx; ++x) {
            const size_t i = (size_t)y * nx + x;
            double p[3], force[3];
            cell_phi(phi, n, i, p);
            const double m = mu[i];
            const double psi0 = psi(P, 2, m);
            const double dtr[3] = {2.0 * p[1] * p[2], 2.0 * p[0] * p[2], 2.0 * p[0] * p[1]};
            double dh[3][3];
            for (int q = 0; q < 3; ++q)
                for (int a = 0; a < 3; ++a) dh[q][a] = dtr[a] + (q == a ? 6.0 * p[q] * (1.0 - p[q]) : 0.0);
            for (int a = 0; a < 3; ++a) {
                force[a] = obs * (p[(a + 1) % 3] + p[(a + 2) % 3]) + tr * p[(a + 1) % 3] * p[(a + 2) % 3];
                for (int b = 0; b < 3; ++b) force[a] += (psi(P, b, m) - psi0) * dh[b][a] / P->Vm;
            }
            for (int dir = 0; dir < 4; ++dir) {
                if ((dir == 2 && y == 0) || (dir == 3 && y == ny - 1)) continue;
                size_t j;
                if (dir == 0) j = (size_t)y * nx + (x == 0 ? nx - 1 : x - 1);
                else if (dir == 1) j = (size_t)y * nx + (x == nx - 1 ? 0 : x + 1);
                else if (dir == 2) j = (size_t)(y - 1) * nx + x;
                else j = (size_t)(y + 1) * nx + x;
                double q[3];
                cell_phi(phi, n, j, q);
                const double dist = dir < 2 ? P->dx : P->dy;
                for (int a = 0; a < 3; ++a)
                    for (int b = a + 1; b < 3; ++b) {
                        const double w = 2 * eps * sigma * (p[a] * q[b] - q[a] * p[b]) / (dist * dist);
                        force[a] += w * q[b];
                        force[b] -= w * q[a];
                    }
            }
            const double mean = (force[0] + force[1] + force[2]) / 3, scale = dt / (local_tau(p, P->tau) * eps);
            double z[3];
            for (int a = 0; a < 3; ++a) z[a] = p[a] - scale * (force[a] - mean);
            int finite = 1;
            for (int a = 0; a < 3; ++a) finite = finite && isfinite(z[a]);
            if (!finite) {
                change = INFINITY;
                for (int a = 0; a < 3; ++a) phi_out[a * n + i] = p[a];
                continue;
            }
            double zp[3];
            if (project_simplex(z, zp) != ORC_OK) {
                *status = ORC_ERR_PROJECTION;
                return change;
            }
            for (int a = 0; a < 3; ++a) change = dmax(change, fabs(zp[a] - p[a]));
            for (int a = 0; a < 3; ++a) phi_out[a * n + i] = zp[a];
        }
    return change;
}

/* ------------------------------------------------------------------ */
/* implicit_diffusion :1010-1094                                       */
/* ------------------------------------------------------------------ */
typedef struct {
    int nx, ny;
    const double* K;
    double dx2, dy2;
} faces_t;

/* weights_at :1022-1028 */
static void weights_at(const faces_t* F, int x, int y, double* w) {
    const int nx = F->nx, ny = F->ny;
    const size_t i = (size_t)y * nx + x;
    const size_t iw = (size_t)y * nx + (x == 0 ? nx - 1 : x - 1);
    const size_t ie = (size_t)y * nx + (x == nx - 1 ? 0 : x + 1);
    const size_t is = (size_t)(y == 0 ? 0 : y - 1) * nx + x;
    const size_t in = (size_t)(y == ny - 1 ? ny - 1 : y + 1) * nx + x;
    const double* K = F->K;
    w[0] = .5 * (K[i] + K[iw]) / F->dx2;
    w[1] = .5 * (K[i] + K[ie]) / F->dx2;
    w[2] = .5 * (K[i] + K[is]) / F->dy2;
    w[3] = .5 * (K[i] + K[in]) / F->dy2;
    if (y == 0) w[2] = 0;
    if (y == ny - 1) w[3] = 0;
}

static void apply_op(const faces_t* F, const double* chi, double dt, const double* in, double* out) {
    const int nx = F->nx, ny = F->ny;
    for (int y = 0; y < ny; ++y)
        for (int x = 0; x < nx; ++x) {
            const size_t i = (size_t)y * nx + x;
            const size_t iw = (size_t)y * nx + (x == 0 ? nx - 1 : x - 1);
            const size_t ie = (size_t)y * nx + (x == nx - 1 ? 0 : x + 1);
            const size_t is = (size_t)(y == 0 ? 0 : y - 1) * nx + x;
            const size_t in_ = (size_t)(y == ny - 1 ? ny - 1 : y + 1) * nx + x;
            double w[4];
            weights_at(F, x, y, w);
            out[i] = chi[i] * in[i] + dt * (w[0] * (in[i] - in[iw]) + w[1] * (in[i] - in[ie]) +
                                            w[2] * (in[i] - in[is]) + w[3] * (in[i] - in[in_]));
        }
}

int orc_implicit_diffusion(const orc_params* P, const double* phi_t, const double* c_old,
                           const double* mu_old, double dt, double mu_ref, double* c_out,
                           double* mu_out, int* ok, int* it_out, double* residual_out,
                           double* closure_out) {
    const int nx = P->nx, ny = P->ny;
    const size_t n = (size_t)nx * ny;
    double* mem = (double*)malloc(11 * n * sizeof(double));
    if (!mem) return ORC_ERR_ALLOC;
    double *chi = mem, *icpt = mem + n, *K = mem + 2 * n, *rhs = mem + 3 * n, *x = mem + 4 * n;
    double *diag = mem + 5 * n, *r = mem + 6 * n, *z = mem + 7 * n, *p = mem + 8 * n, *Ap = mem + 9 * n;
    double* tmp = mem + 10 * n;
    *ok = 0;
    *it_out = 0;
    *residual_out = 0;
    *closure_out = 0;
    for (size_t i = 0; i < n; ++i) {
        double ph[3];
        mix_t m;
        cell_phi(phi_t, n, i, ph);
        if (mixture(P, ph, &m) != ORC_OK) {
            free(mem);
            return ORC_ERR_MIXTURE;
        }
        chi[i] = m.s;
        icpt[i] = composition(mu_ref, &m);
        K[i] = m.K;
        rhs[i] = c_old[i] - icpt[i];
        x[i] = mu_old[i] - mu_ref;
    }
    faces_t F = {nx, ny, K, P->dx * P->dx, P->dy * P->dy};
    apply_op(&F, chi, dt, x, Ap);
    for (int yy = 0; yy < ny; ++yy)
        for (int xx = 0; xx < nx; ++xx) {
            const size_t i = (size_t)yy * nx + xx;
            double w[4];
            weights_at(&F, xx, yy, w);
            diag[i] = chi[i] + dt * (w[0] + w[1] + w[2] + w[3]);
            r[i] = rhs[i] - Ap[i];
            z[i] = r[i] / diag[i];
            p[i] = z[i];
        }
    const double tolerance = P->cg_atol + P->cg_rtol * sqrt(dot(P, rhs, rhs, n, tmp));
    double rz = dot(P, r, z, n, tmp);
    int it;
    int good = 0;
    double residual = 0;
    for (it = 0; it <= P->cg_max; ++it) {
        *it_out = it;
        residual = sqrt(dot(P, r, r, n, tmp));
        *residual_out = residual;
        if (!isfinite(residual)) { free(mem); return ORC_OK; }
        if (residual <= tolerance) {
            apply_op(&F, chi, dt, x, Ap);
            for (size_t i = 0; i < n; ++i) r[i] = rhs[i] - Ap[i];
            residual = sqrt(dot(P, r, r, n, tmp));
            *residual_out = residual;
            if (residual <= tolerance * 1.05) {
                good = 1;
                break;
            }
            for (size_t i = 0; i < n; ++i) {
                z[i] = r[i] / diag[i];
                p[i] = z[i];
            }
            rz = dot(P, r, z, n, tmp);
            continue;
        }
        if (it == P->cg_max) break;
        apply_op(&F, chi, dt, p, Ap);
        const double pAp = dot(P, p, Ap, n, tmp);
        if (!(pAp > 0) || !isfinite(pAp) || !isfinite(rz)) { free(mem); return ORC_OK; }
        const double step = rz / pAp;
        for (size_t i = 0; i < n; ++i) {
            x[i] += step * p[i];
            r[i] -= step * Ap[i];
            z[i] = r[i] / diag[i];
        }
        const double next = dot(P, r, z, n, tmp), beta = next / rz;
        rz = next;
        for (size_t i = 0; i < n; ++i) p[i] = z[i] + beta * p[i];
    }
    if (!good) { free(mem); return ORC_OK; }
    double closure = 0;
    for (int yy = 0; yy < ny; ++yy)
        for (int xx = 0; xx < nx; ++xx) {
            const size_t i = (size_t)yy * nx + xx;
            const size_t iw = (size_t)yy * nx + (xx == 0 ? nx - 1 : xx - 1);
            const size_t ie = (size_t)yy * nx + (xx == nx - 1 ? 0 : xx + 1);
            const size_t is = (size_t)(yy == 0 ? 0 : yy - 1) * nx + xx;
            const size_t in_ = (size_t)(yy == ny - 1 ? ny - 1 : yy + 1) * nx + xx;
            double w[4];
            weights_at(&F, xx, yy, w);
            const double change = dt * (w[0] * (x[iw] - x[i]) + w[1] * (x[ie] - x[i]) + w[2] * (x[is] - x[i]) +
                                        w[3] * (x[in_] - x[i]));
            c_out[i] = c_old[i] + change;
            closure = dmax(closure, fabs(c_out[i] - (icpt[i] + chi[i] * x[i])));
            mu_out[i] = mu_ref + (c_out[i] - icpt[i]) / chi[i];
        }
    *closure_out = closure;
    *ok = closure <= P->closure_tol;
    free(mem);
    return ORC_OK;
}

/* ------------------------------------------------------------------ */
/* valid_state, total_carbon, reduced_energy                          */
/* ------------------------------------------------------------------ */
int orc_valid_state(const orc_params* P, const double* phi, const double* c, const double* mu) {
    const size_t n = (size_t)P->nx * P->ny;
    int good = 1;
    for (size_t i = 0; i < n; ++i) {
        double p[3], sum = 0;
        cell_phi(phi, n, i, p);
        for (int a = 0; a < 3; ++a) {
            good = good && isfinite(p[a]) && p[a] >= -1e-12 && p[a] <= 1 + 1e-12;
            sum += p[a];
        }
        good = good && fabs(sum - 1) < 1e-10;
        good = good && isfinite(mu[i]) && isfinite(c[i]);
        if (!P->phase_only) {
            good = good && c[i] >= -1e-10 && c[i] <= 1 + 1e-10;
            for (int a = 0; a < 3; ++a)
                if (p[a] > .01) {
                    const double cp = (mu[i] - P->B[a]) / (2.0 * P->A[a]); /* c_at_mu :83 */
                    good = good && cp >= -1e-10 && cp <= 1 + 1e-10;
                }
        }
    }
    return good;
}

double orc_total_carbon(const orc_params* P, const double* c) {
    const size_t n = (size_t)P->nx * P->ny;
    return orc_sum(P, c, n) * P->dx * P->dy;
}

/* Per-cell local energy in the blocked order: the cell term, then its east
 * and north face terms, in the reference's order. */
double orc_reduced_energy(const orc_params* P, const double* phi, const double* c, double mu_ref,
                          int* status) {
    const int nx = P->nx, ny = P->ny;
    const size_t n = (size_t)nx * ny;
    const double eps = P->eps, sigma = P->sigma;
    const double obstacle = 16 * sigma / (pi_value() * pi_value() * eps), triple = P->triple / eps;
    const double psiv[3] = {psi(P, 0, mu_ref) - psi(P, 2, mu_ref), psi(P, 1, mu_ref) - psi(P, 2, mu_ref), 0};
    const int seq = P->red_mode != ORC_RED_BLOCKED;
    double* cellv = seq ? NULL : (double*)malloc(n * sizeof(double));
    double total = 0;
    *status = ORC_OK;
    for (int y = 0; y < ny; ++y)
        for (int x = 0; x < nx; ++x) {
            const size_t i = (size_t)y * nx + x;
            double p[3], h[3];
            cell_phi(phi, n, i, p);
            weights(p, h);
            double chem = 0;
            for (int a = 0; a < 3; ++a) chem += h[a] * psiv[a];
            if (!P->phase_only) {
                mix_t m;
                if (mixture(P, p, &m) != ORC_OK) {
                    *status = ORC_ERR_MIXTURE;
                    free(cellv);
                    return NAN;
                }
                const double dc = c[i] - composition(mu_ref, &m);
                chem += .5 * dc * dc / m.s;
            }
            double v = 0;
            const double cellterm = chem / P->Vm + obstacle * (p[0] * p[1] + p[0] * p[2] + p[1] * p[2]) + triple * p[0] * p[1] * p[2];
            if (seq) total += cellterm;
            else v = cellterm;
            for (int dir = 0; dir < 2; ++dir) {
                if (dir == 1 && y == ny - 1) continue;
                const size_t j = dir == 0 ? (size_t)y * nx + (x == nx - 1 ? 0 : x + 1) : (size_t)(y + 1) * nx + x;
                double q[3];
                cell_phi(phi, n, j, q);
                const double dist = dir == 0 ? P->dx : P->dy;
                for (int a = 0; a < 3; ++a)
                    for (int b = a + 1; b < 3; ++b) {
                        const double zz = (p[a] * q[b] - q[a] * p[b]) / dist;
                        if (seq) total += eps * sigma * zz * zz;
                        else v += eps * sigma * zz * zz;
                    }
            }
            if (!seq) cellv[i] = v;
        }
    if (!seq) {
        total = blocked_sum(cellv, n, P->nx, P->red_k);
        free(cellv);
    }
    return total * P->dx * P->dy;
}

/* ------------------------------------------------------------------ */
/* advance :1119-1174                                                  */
/* ------------------------------------------------------------------ */
int orc_advance(const orc_params* P, orc_state* s, double mu_ref, double stop_time, orc_stepinfo* info) {
    const size_t n = (size_t)P->nx * P->ny;
    double dt = s->dt;
    if (P->dt_max < dt) dt = P->dt_max;
    if (stop_time - s->time < dt) dt = stop_time - s->time;
    memset(info, 0, sizeof(*info));
    if (!(dt > 0) || !isfinite(dt)) return ORC_ERR_INVALID_TIME;
    int st = ORC_OK;
    const double e0 = P->check_energy ? orc_reduced_energy(P, s->phi, s->carbon, mu_ref, &st) : 0;
    if (st != ORC_OK) return st;
    const double mass0 = P->phase_only ? 0 : orc_total_carbon(P, s->carbon);
    double* tphi = (double*)malloc(5 * n * sizeof(double));
    if (!tphi) return ORC_ERR_ALLOC;
    double* tc = tphi + 3 * n;
    double* tmu = tphi + 4 * n;
    int why = ORC_WHY_PHASE;
    for (int attempt = 0; attempt <= P->max_retries; ++attempt) {
        /* State trial = s */
        memcpy(tc, s->carbon, n * sizeof(double));
        memcpy(tmu, s->mu, n * sizeof(double));
        orc_stepinfo out;
        memset(&out, 0, sizeof(out));
        out.retries = attempt;
        out.accepted_dt = dt;
        out.max_phase_change = orc_propose_phi(P, s->phi, s->mu, dt, tphi, &st);
        if (st != ORC_OK) { free(tphi); return st; }
        int good = isfinite(out.max_phase_change) && out.max_phase_change <= P->max_phase_change;
        why = ORC_WHY_PHASE;
        if (good) {
            if (P->phase_only) {
                for (size_t i = 0; i < n; ++i) {
                    double p[3];
                    mix_t m;
                    cell_phi(tphi, n, i, p);
                    if (mixture(P, p, &m) != ORC_OK) { free(tphi); return ORC_ERR_MIXTURE; }
                    tc[i] = composition(tmu[i], &m);
                }
            } else {
                int ok = 0, it = 0;
                double res = 0, clo = 0;
                st = orc_implicit_diffusion(P, tphi, s->carbon, s->mu, dt, mu_ref, tc, tmu, &ok, &it, &res, &clo);
                if (st != ORC_OK) { free(tphi); return st; }
                good = ok;
                out.cg_iterations = it;
                out.linear_residual = res;
                out.closure = clo;
                why = ORC_WHY_LINEAR;
            }
        }
        if (good) {
            good = orc_valid_state(P, tphi, tc, tmu);
            why = ORC_WHY_BOUNDS;
        }
        if (good && !P->phase_only) {
            const double m = orc_total_carbon(P, tc);
            good = fabs(m - mass0) <= 1e-11 * dmax(fabs(mass0), P->nx * P->ny * P->dx * P->dy * 1e-6);
            why = ORC_WHY_MASS;
        }
        if (good && P->check_energy) {
            const double e1 = orc_reduced_energy(P, tphi, tc, mu_ref, &st);
            if (st != ORC_OK) { free(tphi); return st; }
            out.energy_change = e1 - e0;
            const double tol = P->energy_atol + P->energy_rtol * dmax(fabs(e0), P->sigma * dmax(P->nx * P->dx, P->ny * P->dy));
            good = isfinite(e1) && e1 <= e0 + tol;
            why = ORC_WHY_ENERGY;
        }
        if (good) {
            memcpy(s->phi, tphi, 3 * n * sizeof(double));
            memcpy(s->carbon, tc, n * sizeof(double));
            memcpy(s->mu, tmu, n * sizeof(double));
            s->time = s->time + dt;
            s->step = s->step + 1;
            s->rejected = s->rejected + attempt;
            s->dt = dmin(P->dt_max, dt * 1.1);
            out.why = -1;
            *info = out;
            free(tphi);
            return ORC_OK;
        }
        dt *= .5;
        if (dt < P->dt_min) {
            info->why = why;
            free(tphi);
            return ORC_ERR_DT_MIN;
        }
    }
    info->why = why;
    free(tphi);
    return ORC_ERR_RETRY_LIMIT;
}

/* ------------------------------------------------------------------ */
/* close_mu :789-792, maybe_shift :2053-2071, shift_rows :2023-2052     */
/* ------------------------------------------------------------------ */
int orc_close_mu(const orc_params* P, orc_state* s) {
    const size_t n = (size_t)P->nx * P->ny;
    for (size_t i = 0; i < n; ++i) {
        double p[3];
        mix_t m;
        cell_phi(s->phi, n, i, p);
        if (mixture(P, p, &m) != ORC_OK) return ORC_ERR_MIXTURE;
        if (!isfinite(s->carbon[i])) return ORC_ERR_NONFINITE_CARBON;
        s->mu[i] = (s->carbon[i] - m.c0) / m.s;
    }
    return ORC_OK;
}

int orc_maybe_shift(const orc_params* P, orc_state* s, int moving_window, double trigger,
                    double target, double c_gamma, int* shifted, double* front_out) {
    const int nx = P->nx, ny = P->ny;
    const size_t n = (size_t)nx * ny;
    *shifted = 0;
    if (!moving_window) return ORC_OK;
    const double* g = s->phi + 2 * n;
    double local = -1;
    for (int y = 0; y < ny - 1; ++y)
        for (int x = 0; x < nx; ++x) {
            const double a = 1 - 2 * g[(size_t)y * nx + x], b = 1 - 2 * g[(size_t)(y + 1) * nx + x];
            if (a >= 0 && b < 0) local = dmax(local, ((y + 1) - .5 + a / (a - b)) * P->dy);
        }
    if (front_out) *front_out = local;
    if (local < trigger) return ORC_OK;
    const int rows = (int)floor((local - target) / P->dy);
    double gamma = 0;
    for (int y = 0; y < ny; ++y)
        if (y < rows)
            for (int x = 0; x < nx; ++x) gamma = dmax(gamma, g[(size_t)y * nx + x]);
    if (gamma > 1e-6) return ORC_ERR_UNSAFE_SHIFT;
    /* shift_rows */
    if (rows <= 0 || rows >= ny || c_gamma < 0 || c_gamma > 1) return ORC_ERR_INVALID_SHIFT;
    const double added = rows * nx * c_gamma * P->dx * P->dy;
    double removed;
    if (P->red_mode == ORC_RED_BLOCKED) {
        const size_t m = (size_t)rows * nx;
        double* tmp = (double*)malloc(m * sizeof(double));
        for (size_t i = 0; i < m; ++i) tmp[i] = s->carbon[i] * P->dx * P->dy;
        removed = blocked_sum(tmp, m, P->nx, P->red_k);
        free(tmp);
    } else {
        removed = 0;
        for (int y = 0; y < rows; ++y)
            for (int x = 0; x < nx; ++x) removed += s->carbon[(size_t)y * nx + x] * P->dx * P->dy;
    }
    for (int y = 0; y < ny; ++y)
        for (int x = 0; x < nx; ++x) {
            const size_t i = (size_t)y * nx + x;
            if (y + rows < ny) {
                const size_t j = (size_t)(y + rows) * nx + x;
                s->carbon[i] = s->carbon[j];
                for (int a = 0; a < 3; ++a) s->phi[a * n + i] = s->phi[a * n + j];
            } else {
                s->carbon[i] = c_gamma;
                s->phi[i] = 0;
                s->phi[n + i] = 0;
                s->phi[2 * n + i] = 1;
            }
        }
    const int st = orc_close_mu(P, s);
    if (st != ORC_OK) return st;
    s->shift_distance += rows * P->dy;
    s->shifts++;
    s->carbon_added += added;
    s->carbon_removed += removed;
    *shifted = 1;
    return ORC_OK;
}

/* ------------------------------------------------------------------ */
/* smooth_interface_only :869-958 with interface_energy_only :822-841  */
/* ------------------------------------------------------------------ */
static double interface_energy(const orc_params* P, const double* phi) {
    const int nx = P->nx, ny = P->ny;
    const size_t n = (size_t)nx * ny;
    const double eps = P->eps, sigma = P->sigma;
    const double obstacle = 16 * sigma / (pi_value() * pi_value() * eps), triple = P->triple / eps;
    const int seq = P->red_mode != ORC_RED_BLOCKED;
    double* cellv = seq ? NULL : (double*)malloc(n * sizeof(double));
    double total = 0;
    for (int y = 0; y < ny; ++y)
        for (int x = 0; x < nx; ++x) {
            const size_t i = (size_t)y * nx + x;
            double p[3];
            cell_phi(phi, n, i, p);
            const double cellterm = obstacle * (p[0] * p[1] + p[0] * p[2] + p[1] * p[2]) + triple * p[0] * p[1] * p[2];
            double v = 0;
            if (seq) total += cellterm;
            else v = cellterm;
            for (int dir = 0; dir < 2; ++dir) {
                if (dir == 1 && y == ny - 1) continue;
                const size_t j = dir == 0 ? (size_t)y * nx + (x == nx - 1 ? 0 : x + 1) : (size_t)(y + 1) * nx + x;
                double q[3];
                cell_phi(phi, n, j, q);
                const double dist = dir == 0 ? P->dx : P->dy;
                for (int a = 0; a < 3; ++a)
                    for (int b = a + 1; b < 3; ++b) {
                        const double zz = (p[a] * q[b] - q[a] * p[b]) / dist;
                        if (seq) total += eps * sigma * zz * zz;
                        else v += eps * sigma * zz * zz;
                    }
            }
            if (!seq) cellv[i] = v;
        }
    if (!seq) {
        total = blocked_sum(cellv, n, P->nx, P->red_k);
        free(cellv);
    }
    return total * P->dx * P->dy;
}

int orc_smooth(const orc_params* P, double* phi, int max_steps, int min_steps, double energy_rtol,
               double max_phase_change, double* report) {
    const int nx = P->nx, ny = P->ny;
    const size_t n = (size_t)nx * ny;
    for (int k = 0; k < 6; ++k) report[k] = 0;
    const double energy_initial = interface_energy(P, phi);
    report[2] = energy_initial;
    double previous = energy_initial;
    const double eps = P->eps, sigma = P->sigma;
    const double obs = 16 * sigma / (pi_value() * pi_value() * eps), tr = P->triple / eps;
    const double base_scale = .12 * dmin(P->dx * P->dx, P->dy * P->dy) / (eps * sigma);
    double* force = (double*)malloc(6 * n * sizeof(double));
    double* trial = force + 3 * n;
    int status = ORC_OK;
    for (int step = 1; step <= max_steps; ++step) {
        double force_max = 0;
        for (int y = 0; y < ny; ++y)
            for (int x = 0; x < nx; ++x) {
                const size_t i = (size_t)y * nx + x;
                double p[3], f[3];
                cell_phi(phi, n, i, p);
                for (int a = 0; a < 3; ++a) f[a] = obs * (p[(a + 1) % 3] + p[(a + 2) % 3]) + tr * p[(a + 1) % 3] * p[(a + 2) % 3];
                for (int dir = 0; dir < 4; ++dir) {
                    if ((dir == 2 && y == 0) || (dir == 3 && y == ny - 1)) continue;
                    size_t j;
                    if (dir == 0) j = (size_t)y * nx + (x == 0 ? nx - 1 : x - 1);
                    else if (dir == 1) j = (size_t)y * nx + (x == nx - 1 ? 0 : x + 1);
                    else if (dir == 2) j = (size_t)(y - 1) * nx + x;
                    else j = (size_t)(y + 1) * nx + x;
                    double q[3];
                    cell_phi(phi, n, j, q);
                    const double dist = dir < 2 ? P->dx : P->dy;
                    for (int a = 0; a < 3; ++a)
                        for (int b = a + 1; b < 3; ++b) {
                            const double w = 2 * eps * sigma * (p[a] * q[b] - q[a] * p[b]) / (dist * dist);
                            f[a] += w * q[b];
                            f[b] -= w * q[a];
                        }
                }
                const double mean = (f[0] + f[1] + f[2]) / 3;
                for (int a = 0; a < 3; ++a) {
                    force[a * n + i] = f[a] - mean;
                    force_max = dmax(force_max, fabs(force[a * n + i]));
                }
            }
        if (!isfinite(force_max)) { status = -1; break; }
        if (force_max == 0) {
            report[0] = step - 1;
            report[1] = 1;
            report[4] = 0;
            report[5] = 0;
            break;
        }
        double scale = dmin(base_scale, max_phase_change / force_max);
        int accepted = 0;
        double new_energy = previous, actual_change = 0;
        for (int bt = 0; bt < 24 && !accepted; ++bt) {
            actual_change = 0;
            for (size_t i = 0; i < n; ++i) {
                double p[3], z[3], zp[3];
                cell_phi(phi, n, i, p);
                for (int a = 0; a < 3; ++a) z[a] = p[a] - scale * force[a * n + i];
                if (!isfinite(z[0]) || !isfinite(z[1]) || !isfinite(z[2]) || project_simplex(z, zp) != ORC_OK) {
                    free(force);
                    return ORC_ERR_PROJECTION;
                }
                for (int a = 0; a < 3; ++a) actual_change = dmax(actual_change, fabs(zp[a] - p[a]));
                for (int a = 0; a < 3; ++a) trial[a * n + i] = zp[a];
            }
            new_energy = interface_energy(P, trial);
            double m = 1.0;
            if (m < fabs(previous)) m = fabs(previous);
            if (m < fabs(energy_initial)) m = fabs(energy_initial);
            const double etol = 1e-14 * m;
            if (isfinite(new_energy) && new_energy <= previous + etol) accepted = 1;
            else scale *= .5;
        }
        if (!accepted) { status = -2; break; }
        memcpy(phi, trial, 3 * n * sizeof(double));
        const double rel = fabs(previous - new_energy) / dmax(fabs(previous), 1e-300);
        report[0] = step;
        report[3] = new_energy;
        report[4] = rel;
        report[5] = actual_change;
        previous = new_energy;
        if (step >= min_steps && rel < energy_rtol && actual_change < 1e-5) {
            report[1] = 1;
            break;
        }
    }
    if (report[0] == 0) report[3] = energy_initial;
    free(force);
    return status;
}
