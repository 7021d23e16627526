/*
 * oracle.c - CPU restatement of the reference DBA time step (TEST INFRASTRUCTURE).
 *
 * Not the product: only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg load it. See oracle.h for the reference files it follows.
 *
 * Built with -O2 -ffp-contract=off so every +,-,*,/ is one IEEE operation in
 * source order, like the HIP kernels (also built with -ffp-contract=off).
 *
 * Parity pinning status per function (the checks are in tests/test_oracle_*.py):
 *   ch_*      pinned: 5/9-point Laplacian Fourier symbols, exact discrete mean law,
 *             single-mode linear growth factor, conserved-DBA total, fixed points.
 *   k93_*     pinned: fixed points, frozen-phi heat-equation modes, hand-computed
 *             single-cell update, anisotropy vs atan2/cos/sin, D4 symmetry.
 *   kh_*      pinned: bit-exact against the compiled reference CPU program
 *             (KH_CPU.cpp) on two cases; uniform fixed point; conservation with
 *             periodic and wall boundaries; transpose symmetry; 2nd-order acoustics.
 *   mb_*      pinned: bit-exact against the compiled reference CPU program
 *             (base_CPU.cpp) on an 8^3 case and a 2D plane; uniform fixed point.
 *   k94_*     pinned to float32 resolution against the compiled (float32)
 *             reference CPU program on the central block.
 *   DBA modes (all models): always-awake equals regular bitwise, a tile that
 *             never fires stays frozen bitwise, conserved totals invariant, and
 *             the stale-neighbour data flow pinned on tiny grids with prescribed
 *             awake sets against numpy restatements of the reference's kernel
 *             sequences (tests/dba_sequence.py, tests/test_oracle_stale.py):
 *             stale mu (CH, Model B), dropped latent heat in p-awake/T-asleep
 *             tiles (K93, K94), stale primitives/gradients/primes and face
 *             fluxes (KH), the mean shift on every cell (conserved CH, Model B,
 *             KH). Each of those tests also builds this file with one planted
 *             mistake (ORA_MUTANT below) and checks that the mistake is caught.
 *   dilate / compact  pinned: brute force over all small flag patterns.
 */
#define _GNU_SOURCE
#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#include "oracle.h"

#define MAXF 4
#define MAXW 32

/* Planted mistakes for the tests that pin the DBA data flow
   (tests/test_oracle_stale.py compiles copies with -DORA_MUTANT=k and checks that
   each is caught). The default build, the only one anything else loads, has none.
     1  chemical potential recomputed on every cell every step (fresh, not the
        value stored when a sleeping tile last ran) - CH, Model B
     2  latent heat applied and kept in p-awake / T-asleep tiles - K93, K94
     3  KH primitives, gradients and primes recomputed on every cell every step
     4  KH face fluxes recomputed on every cell every step
     5  conserved mean shift spread over the awake cells only - CH, Model B, KH */
#ifndef ORA_MUTANT
#define ORA_MUTANT 0
#endif

struct ora_sim {
    dba_params p;
    int nx, ny, nz;
    long n;
    int tx, ty, tz, ntx, nty, ntz, nt;
    double* f[MAXF];    /* current fields */
    double* fn[MAXF];   /* new fields */
    double* w[MAXW];    /* derivative / derived arrays */
    uint8_t* awake[2];
    uint8_t* next[2];
    int32_t* hist;
    int hist_n, hist_cap;
    long steps;
    double time, last_dt;
    double dt_kh;        /* KH: time step of the current step (fixed or CFL) */
    /* derived constants (identical expressions in the HIP host code) */
    double invdx2, inv6dx2, inv2dx, dt_tau, alpha_pi, cos_jt0, sin_jt0, k_epsp;
};

/* ------------------------------------------------------------------ helpers */

static inline int map_i(int i, int n, int bc) {
    if (bc == DBA_BC_PERIODIC) {
        i %= n;
        if (i < 0) i += n;
        return i;
    }
    if (i < 0) return -1 - i;
    if (i >= n) return 2 * n - 1 - i;
    return i;
}

static inline long idx2(const ora_sim* s, int x, int y) {
    return (long)map_i(x, s->nx, s->p.bc_x) + (long)map_i(y, s->ny, s->p.bc_y) * s->nx;
}

static inline int tile_of(const ora_sim* s, int x, int y, int z) {
    return x / s->tx + (y / s->ty) * s->ntx + (z / s->tz) * s->ntx * s->nty;
}

static inline int active(const ora_sim* s, int set, int x, int y, int z) {
    if (s->p.mode == DBA_MODE_REGULAR) return 1;
    return s->awake[set][tile_of(s, x, y, z)];
}

/* 5-point and 9-point (isotropic, the 27-point stencil of
   example_Reactive_ModelB/kernel.cu:290-347 reduced to one z-plane) Laplacians. */
static inline double lap5(const ora_sim* s, const double* f, int x, int y) {
    double s1 = (f[idx2(s, x + 1, y)] + f[idx2(s, x - 1, y)]) + (f[idx2(s, x, y + 1)] + f[idx2(s, x, y - 1)]);
    return (s1 - 4.0 * f[idx2(s, x, y)]) * s->invdx2;
}
static inline double lap9(const ora_sim* s, const double* f, int x, int y) {
    double s1 = (f[idx2(s, x + 1, y)] + f[idx2(s, x - 1, y)]) + (f[idx2(s, x, y + 1)] + f[idx2(s, x, y - 1)]);
    double s2 = (f[idx2(s, x + 1, y + 1)] + f[idx2(s, x - 1, y + 1)]) + (f[idx2(s, x + 1, y - 1)] + f[idx2(s, x - 1, y - 1)]);
    return ((4.0 * s1 + s2) - 20.0 * f[idx2(s, x, y)]) * s->inv6dx2;
}

void ora_dilate(const uint8_t* N, uint8_t* A, int ntx, int nty, int ntz, int kind,
                int wrap_x, int wrap_y, int wrap_z) {
    /* example_Dendritic_growth/kernel.cu:577-634 (faces, periodic),
       example_Kelvin_Helmholtz/KH_DBA.cu:988-1044 (3x3 incl. diagonals + z faces). */
    long nt = (long)ntx * nty * ntz;
    memset(A, 0, (size_t)nt);
    for (int k = 0; k < ntz; k++)
        for (int j = 0; j < nty; j++)
            for (int i = 0; i < ntx; i++) {
                long t = i + (long)j * ntx + (long)k * ntx * nty;
                if (!N[t]) continue;
                for (int dk = -1; dk <= 1; dk++)
                    for (int dj = -1; dj <= 1; dj++)
                        for (int di = -1; di <= 1; di++) {
                            int nz_ = (di != 0) + (dj != 0) + (dk != 0);
                            if (kind == 0 && nz_ > 1) continue;
                            if (kind == 1 && dk != 0 && (di != 0 || dj != 0)) continue;
                            int ii = i + di, jj = j + dj, kk = k + dk;
                            if (ii < 0 || ii >= ntx) { if (!wrap_x) continue; ii = (ii + ntx) % ntx; }
                            if (jj < 0 || jj >= nty) { if (!wrap_y) continue; jj = (jj + nty) % nty; }
                            if (kk < 0 || kk >= ntz) { if (!wrap_z) continue; kk = (kk + ntz) % ntz; }
                            A[ii + (long)jj * ntx + (long)kk * ntx * nty] = 1;
                        }
            }
}

int ora_compact(const uint8_t* A, int nt, int32_t* out) {
    int c = 0;
    for (int t = 0; t < nt; t++)
        if (A[t]) out[c++] = t;
    return c;
}

static void record_history(ora_sim* s) {
    if (s->hist_n >= s->hist_cap) return;
    int c0 = 0, c1 = 0;
    if (s->p.mode == DBA_MODE_REGULAR) {
        c0 = c1 = s->nt;
    } else {
        for (int t = 0; t < s->nt; t++) { c0 += s->awake[0][t]; c1 += s->awake[1][t]; }
    }
    s->hist[2 * s->hist_n] = c0;
    s->hist[2 * s->hist_n + 1] = c1;
    s->hist_n++;
}

/* Wakeup_Next + Wakeup_Neighbor for one set, given per-cell criterion flags
   already folded into next[] (example_Reactive_ModelB/kernel.cu:657-763). */
static void dilate_set(ora_sim* s, int set) {
    int wx = s->p.bc_x == DBA_BC_PERIODIC, wy = s->p.bc_y == DBA_BC_PERIODIC, wz = s->p.bc_z == DBA_BC_PERIODIC;
    int kind = s->p.model == DBA_MODEL_KH ? 1 : 0;
    ora_dilate(s->next[set], s->awake[set], s->ntx, s->nty, s->ntz, kind, wx, wy, wz);
}

/* --------------------------------------------------------------- CH model */
/* dc/dt = M lap(c^3 - c - kappa lap c) - Gamma (c - cbar), 2D.
   Loop structure of example_Reactive_ModelB/CPU_version/base_CPU.cpp:563-596:
   Calc_mu (all cells) -> Update -> Swap; DBA semantics of
   example_Reactive_ModelB/kernel.cu:1094-1131 (mu of sleeping blocks stays stale).
   w[0] = mu, w[1] = metric |lap c|, w[2] = change (conserved). */

static inline double ch_lap(const ora_sim* s, const double* f, int x, int y) {
    return s->p.stencil == 5 ? lap5(s, f, x, y) : lap9(s, f, x, y);
}

static void ch_calc_mu(ora_sim* s, int y0, int y1) {
    const double* c = s->f[0];
    double* mu = s->w[0];
    double* metric = s->w[1];
    for (int y = y0; y < y1; y++)
        for (int x = 0; x < s->nx; x++) {
            if (!active(s, 0, x, y, 0) && ORA_MUTANT != 1) continue;
            long i = (long)x + (long)y * s->nx;
            double lc = ch_lap(s, c, x, y);
            double ci = c[i];
            mu[i] = (ci * ci * ci - ci) - s->p.kappa * lc;
            metric[i] = fabs(lc);
        }
}

static void ch_update(ora_sim* s, int y0, int y1) {
    const double* c = s->f[0];
    double* cn = s->fn[0];
    const double* mu = s->w[0];
    double* chg = s->w[2];
    int cons = s->p.mode == DBA_MODE_CONSERVED;
    for (int y = y0; y < y1; y++)
        for (int x = 0; x < s->nx; x++) {
            long i = (long)x + (long)y * s->nx;
            if (!active(s, 0, x, y, 0)) {
                if (cons) chg[i] = 0.0;  /* example_Reactive_ModelB/Conserved/kernel.cu:688-700 */
                continue;
            }
            double lm = ch_lap(s, mu, x, y);
            double change = s->p.dt * (s->p.mobility * lm - s->p.react_gamma * (c[i] - s->p.cbar));
            cn[i] = c[i] + change;
            if (cons) chg[i] = change;
        }
}

static void ch_step(ora_sim* s) {
    long n = s->n;
    ch_calc_mu(s, 0, s->ny);
    ch_update(s, 0, s->ny);
    if (s->p.mode == DBA_MODE_REGULAR) {
        memcpy(s->f[0], s->fn[0], sizeof(double) * n);
        return;
    }
    if (s->p.mode == DBA_MODE_CONSERVED) {
        /* thrust::reduce of change / N_total, subtracted from every cell, then a
           full swap (example_Reactive_ModelB/Conserved/kernel.cu:1152-1173). */
        double sum = 0.0;
        for (long i = 0; i < n; i++) sum += s->w[2][i];
#if ORA_MUTANT == 5
        long na = 0;
        for (int y = 0; y < s->ny; y++)
            for (int x = 0; x < s->nx; x++) na += active(s, 0, x, y, 0);
        for (int y = 0; y < s->ny; y++)
            for (int x = 0; x < s->nx; x++) {
                long i = (long)x + (long)y * s->nx;
                if (active(s, 0, x, y, 0)) s->f[0][i] = s->fn[0][i] - sum / (double)na;
            }
#else
        double avg = sum / (double)n;
        for (long i = 0; i < n; i++) s->fn[0][i] -= avg;
        memcpy(s->f[0], s->fn[0], sizeof(double) * n);
#endif
    } else {
        for (int y = 0; y < s->ny; y++)
            for (int x = 0; x < s->nx; x++)
                if (active(s, 0, x, y, 0)) {
                    long i = (long)x + (long)y * s->nx;
                    s->f[0][i] = s->fn[0][i];
                }
    }
    /* Wakeup_Next: any(metric > Metric_eps) over awake blocks */
    memset(s->next[0], 0, (size_t)s->nt);
    for (int y = 0; y < s->ny; y++)
        for (int x = 0; x < s->nx; x++) {
            int t = tile_of(s, x, y, 0);
            if (s->awake[0][t] && s->w[1][(long)x + (long)y * s->nx] > s->p.thresh_a) s->next[0][t] = 1;
        }
    dilate_set(s, 0);
}

/* -------------------------------------------------------- Kobayashi-93 model */
/* tau dphi/dt = -d/dx(eps eps' phi_y) + d/dy(eps eps' phi_x) + eps^2 lap phi
                 + phi(1-phi)(phi - 1/2 + m),  m = (alpha/pi) atan(gamma (T_eq - T)),
   dT/dt = lap T + K dphi/dt,  eps = eps_bar (1 + delta cos j(theta - theta0)),
   theta = angle of grad phi. The angle enters only through cos/sin of j*theta,
   evaluated as the j-th power of (phi_x + i phi_y)/|grad phi| (theta = 0 where
   grad phi = 0, as atan2(0,0) = 0).
   DBA semantics follow the reference dendrite (example_Dendritic_growth/kernel.cu:322-634):
   separate p and T sets, T update K dphi only where the block is p-awake, latent
   heat dropped in p-awake/T-asleep blocks.
   w[0] = Fx = eps eps' phi_x, w[1] = Fy (both on a 1-cell ghost frame),
   w[2] = eps^2, w[3] = lap phi, w[4] = lap T, w[5] = metric_T. */

static inline void k93_aniso(const ora_sim* s, double gx, double gy, double* eps, double* a) {
    double g2 = gx * gx + gy * gy;
    double cj = 1.0, sj = 0.0;
    if (g2 != 0.0) {
        double r = sqrt(g2);
        double u = gx / r, v = gy / r;
        double cr = u, ci = v;
        for (int k = 1; k < s->p.aniso_j; k++) {
            double t = cr * u - ci * v;
            ci = cr * v + ci * u;
            cr = t;
        }
        cj = cr;
        sj = ci;
    }
    double ca = cj * s->cos_jt0 + sj * s->sin_jt0;
    double sa = sj * s->cos_jt0 - cj * s->sin_jt0;
    double e = s->p.eps_bar * (1.0 + s->p.delta * ca);
    double ep = s->k_epsp * sa;
    *eps = e;
    *a = e * ep;
}

static inline long fidx(const ora_sim* s, int x, int y) { /* ghost-frame index, x,y in [-1, n] */
    return (long)(x + 1) + (long)(y + 1) * (s->nx + 2);
}

static void k93_finite_diff(ora_sim* s, int y0, int y1) {
    const double* p = s->f[0];
    const double* T = s->f[1];
    /* Fx, Fy on rows [y0-1, y1] when the slab touches the frame */
    int ya = y0 == 0 ? -1 : y0, yb = y1 == s->ny ? s->ny + 1 : y1;
    for (int y = ya; y < yb; y++)
        for (int x = -1; x <= s->nx; x++) {
            double gx = (p[idx2(s, x + 1, y)] - p[idx2(s, x - 1, y)]) * s->inv2dx;
            double gy = (p[idx2(s, x, y + 1)] - p[idx2(s, x, y - 1)]) * s->inv2dx;
            double e, a;
            k93_aniso(s, gx, gy, &e, &a);
            long fi = fidx(s, x, y);
            s->w[0][fi] = a * gx;
            s->w[1][fi] = a * gy;
            if (x >= 0 && x < s->nx && y >= 0 && y < s->ny) {
                long i = (long)x + (long)y * s->nx;
                s->w[2][i] = e * e;
                s->w[3][i] = lap5(s, p, x, y);
                s->w[4][i] = lap5(s, T, x, y);
            }
        }
}

static void k93_evolve(ora_sim* s, int y0, int y1) {
    const double* p = s->f[0];
    const double* T = s->f[1];
    double* pn = s->fn[0];
    double* Tn = s->fn[1];
    const double* Fx = s->w[0];
    const double* Fy = s->w[1];
    for (int y = y0; y < y1; y++)
        for (int x = 0; x < s->nx; x++) {
            long i = (long)x + (long)y * s->nx;
            int ap = active(s, 0, x, y, 0), at = active(s, 1, x, y, 0);
            double pi_ = p[i];
            if (ap) {
                double cross = ((Fx[fidx(s, x, y + 1)] - Fx[fidx(s, x, y - 1)]) -
                                (Fy[fidx(s, x + 1, y)] - Fy[fidx(s, x - 1, y)])) * s->inv2dx;
                double m = s->alpha_pi * atan(s->p.gamma * (s->p.t_eq - T[i]));
                double rhs = (cross + s->w[2][i] * s->w[3][i]) + (pi_ * (1.0 - pi_)) * ((pi_ - 0.5) + m);
                pn[i] = pi_ + s->dt_tau * rhs;
            }
            if (at) {
                double t1 = T[i] + s->p.dt * s->w[4][i];
                Tn[i] = ap ? t1 + s->p.latent_k * (pn[i] - pi_) : t1;
                s->w[5][i] = fabs(Tn[i] - T[i]);
            }
#if ORA_MUTANT == 2
            else if (ap) {
                Tn[i] = T[i] + s->p.latent_k * (pn[i] - pi_);
            }
#endif
        }
}

static void k93_step(ora_sim* s) {
    long n = s->n;
    k93_finite_diff(s, 0, s->ny);
    k93_evolve(s, 0, s->ny);
    if (s->p.mode == DBA_MODE_REGULAR) {
        memcpy(s->f[0], s->fn[0], sizeof(double) * n);
        memcpy(s->f[1], s->fn[1], sizeof(double) * n);
        return;
    }
    /* Wakeup_Next (example_Dendritic_growth/kernel.cu:496-574) */
    memset(s->next[0], 0, (size_t)s->nt);
    memset(s->next[1], 0, (size_t)s->nt);
    double mp = s->p.thresh_a, mT = s->p.thresh_b;
    for (int y = 0; y < s->ny; y++)
        for (int x = 0; x < s->nx; x++) {
            long i = (long)x + (long)y * s->nx;
            int t = tile_of(s, x, y, 0);
            if (s->awake[0][t]) {
                s->f[0][i] = s->fn[0][i];
                if (s->fn[0][i] >= mp && s->fn[0][i] <= (1.0 - mp)) s->next[0][t] = 1;
            }
            if (s->awake[1][t]) {
                s->f[1][i] = s->fn[1][i];
                if (s->w[5][i] >= mT) s->next[1][t] = 1;
            }
#if ORA_MUTANT == 2
            else if (s->awake[0][t]) {
                s->f[1][i] = s->fn[1][i];
            }
#endif
        }
    dilate_set(s, 0);
    dilate_set(s, 1);
}


/* ---------------------------------------------------------------- KH model */
/* 2D compressible Euler, MUSCL-Hancock with Rusanov fluxes, in the arithmetic
   of example_Kelvin_Helmholtz/CPU_version/KH_CPU.cpp:263-515 (getPrimitive,
   extrap_in_time, extrapolateInSpaceToFace, getFlux, applyFluxes) and the DBA
   step of example_Kelvin_Helmholtz/KH_DBA.cu:1436-1500 (prims, gradients,
   primes and fluxes of sleeping blocks stay stale and are read as such).
   Generalised: gas gamma, CFL time step (example_Kelvin_Helmholtz/CPU_version/KH.py:241),
   reflective y walls through mirrored ghost cells (normal velocity negated),
   density-gradient detector (example_Kelvin_Helmholtz/KH_DBA.cu:618).
   Primitive, gradient, prime and flux arrays carry a 2-cell ghost frame.
   Field order: U = Mass, Momx, Momy, Energy; prims r, vx, vy, p;
   grads r_x r_y p_x p_y vx_x vx_y vy_x vy_y (reference order). */

enum { KP_R = 0, KP_VX = 1, KP_VY = 2, KP_P = 3 };

static inline long pidx(const ora_sim* s, int x, int y) { /* padded index, x,y in [-2, n+1] */
    return (long)(x + 2) + (long)(y + 2) * (s->nx + 4);
}

/* ghost source: mapped real cell and sign for vy (wall mirror in y) */
static inline void kh_map(const ora_sim* s, int x, int y, int* mx, int* my, double* sgn) {
    *mx = map_i(x, s->nx, s->p.bc_x);
    *my = map_i(y, s->ny, s->p.bc_y == DBA_BC_WALL ? DBA_BC_NOFLUX : s->p.bc_y);
    *sgn = (s->p.bc_y == DBA_BC_WALL && (y < 0 || y >= s->ny)) ? -1.0 : 1.0;
}

static void kh_prims_of(const ora_sim* s, long i, double* r, double* vx, double* vy, double* pr) {
    double vol = s->p.dx * s->p.dy, g = s->p.gas_gamma;
    double rho = s->f[0][i] / vol;
    double u = s->f[1][i] / rho / vol;
    double v = s->f[2][i] / rho / vol;
    *r = rho;
    *vx = u;
    *vy = v;
    *pr = (s->f[3][i] / vol - 0.5 * rho * (u * u + v * v)) * (g - 1.0);
}

static inline double kh_speed(const ora_sim* s, double r, double u, double v, double pr) {
    return sqrt(s->p.gas_gamma * pr / r) + sqrt(u * u + v * v);
}

static void kh_flux(const ora_sim* s, double rL, double rR, double pL, double pR, double uL, double uR, double vL,
                    double vR, double* fm, double* fe, double* fu, double* fv) {
    /* getFlux, example_Kelvin_Helmholtz/KH_DBA.cu:662-697 */
    double g = s->p.gas_gamma;
    double en_L = pL / (g - 1.0) + 0.5 * rL * (uL * uL + vL * vL);
    double en_R = pR / (g - 1.0) + 0.5 * rR * (uR * uR + vR * vR);
    double rho_star = 0.5 * (rL + rR);
    double momx_star = 0.5 * (rL * uL + rR * uR);
    double momy_star = 0.5 * (rL * vL + rR * vR);
    double en_star = 0.5 * (en_L + en_R);
    double p_star = (g - 1.0) * (en_star - 0.5 * (momx_star * momx_star + momy_star * momy_star) / rho_star);
    double flux_Mass = momx_star;
    double flux_Momx = momx_star * momx_star / rho_star + p_star;
    double flux_Momy = momx_star * momy_star / rho_star;
    double flux_Energy = (en_star + p_star) * momx_star / rho_star;
    double C_L = sqrt(g * pL / rL) + fabs(uL);
    double C_R = sqrt(g * pR / rR) + fabs(uR);
    double C = (C_L >= C_R) ? C_L : C_R;
    *fm = flux_Mass - C * 0.5 * (rL - rR);
    *fu = flux_Momx - C * 0.5 * (rL * uL - rR * uR);
    *fv = flux_Momy - C * 0.5 * (rL * vL - rR * vR);
    *fe = flux_Energy - C * 0.5 * (en_L - en_R);
}

/* w[0..3] prims, w[4..11] grads, w[12..15] primes, w[16..19] F (x faces),
   w[20..23] G (y faces), w[24..27] conserved changes, w[28] metric. */
static void kh_grads_at(ora_sim* s, int x, int y) {
    long c = pidx(s, x, y);
    long xm = pidx(s, x - 1, y), xp = pidx(s, x + 1, y), ym = pidx(s, x, y - 1), yp = pidx(s, x, y + 1);
    double dx = s->p.dx, dy = s->p.dy, dt = s->dt_kh;
    double** P = &s->w[0];
    /* Grad_x / Grad_y, example_Kelvin_Helmholtz/KH_DBA.cu:249-275 */
    double r_x = (P[KP_R][xp] - P[KP_R][xm]) / 2.0 / dx, r_y = (P[KP_R][yp] - P[KP_R][ym]) / 2.0 / dy;
    double p_x = (P[KP_P][xp] - P[KP_P][xm]) / 2.0 / dx, p_y = (P[KP_P][yp] - P[KP_P][ym]) / 2.0 / dy;
    double u_x = (P[KP_VX][xp] - P[KP_VX][xm]) / 2.0 / dx, u_y = (P[KP_VX][yp] - P[KP_VX][ym]) / 2.0 / dy;
    double v_x = (P[KP_VY][xp] - P[KP_VY][xm]) / 2.0 / dx, v_y = (P[KP_VY][yp] - P[KP_VY][ym]) / 2.0 / dy;
    double r = P[KP_R][c], u = P[KP_VX][c], v = P[KP_VY][c], pr = P[KP_P][c];
    double g = s->p.gas_gamma;
    s->w[4][c] = r_x; s->w[5][c] = r_y; s->w[6][c] = p_x; s->w[7][c] = p_y;
    s->w[8][c] = u_x; s->w[9][c] = u_y; s->w[10][c] = v_x; s->w[11][c] = v_y;
    /* extrap_in_time, example_Kelvin_Helmholtz/KH_DBA.cu:524-531 */
    s->w[12][c] = r - 0.5 * dt * (u * r_x + r * u_x + v * r_y + r * v_y);
    s->w[13][c] = u - 0.5 * dt * (u * u_x + v * u_y + (1.0 / r) * p_x);
    s->w[14][c] = v - 0.5 * dt * (u * v_x + v * v_y + (1.0 / r) * p_y);
    s->w[15][c] = pr - 0.5 * dt * (g * pr * (u_x + v_y) + u * p_x + v * p_y);
    if (s->p.metric_kind == DBA_KH_METRIC_DENSITY)
        s->w[28][c] = fabs(r_x) + fabs(r_y);
    else
        s->w[28][c] = fabs(r_x) + fabs(r_y) + fabs(p_x) + fabs(p_y) + fabs(u_x) + fabs(u_y) + fabs(v_x) + fabs(v_y);
}

/* x-face flux between (x,y) and (x+1,y) stored at (x,y); y-face between (x,y) and (x,y+1) */
static void kh_face_fluxes(ora_sim* s, int x, int y, int do_x, int do_y) {
    long c = pidx(s, x, y), xp = pidx(s, x + 1, y), yp = pidx(s, x, y + 1);
    double dx = s->p.dx, dy = s->p.dy;
    double* PR[4] = {s->w[12], s->w[13], s->w[14], s->w[15]};            /* r, vx, vy, p primes */
    double* GX[4] = {s->w[4], s->w[8], s->w[10], s->w[6]};               /* r_x, vx_x, vy_x, p_x */
    double* GY[4] = {s->w[5], s->w[9], s->w[11], s->w[7]};
    double f0[4], f1[4], f2[4], f3[4];
    for (int q = 0; q < 4; q++) {
        /* extrapolateInSpaceToFace, example_Kelvin_Helmholtz/KH_DBA.cu:656-659 */
        f0[q] = PR[q][xp] - GX[q][xp] * dx / 2.0;
        f1[q] = PR[q][c] + GX[q][c] * dx / 2.0;
        f2[q] = PR[q][yp] - GY[q][yp] * dy / 2.0;
        f3[q] = PR[q][c] + GY[q][c] * dy / 2.0;
    }
    double fm, fe, fu, fv;
    if (do_x) {
        kh_flux(s, f0[0], f1[0], f0[3], f1[3], f0[1], f1[1], f0[2], f1[2], &fm, &fe, &fu, &fv);
        s->w[16][c] = fm; s->w[17][c] = fu; s->w[18][c] = fv; s->w[19][c] = fe;
    }
    if (do_y) {
        /* momentum roles swapped for the y face, example_Kelvin_Helmholtz/KH_DBA.cu:738-740 */
        kh_flux(s, f2[0], f3[0], f2[3], f3[3], f2[2], f3[2], f2[1], f3[1], &fm, &fe, &fv, &fu);
        s->w[20][c] = fm; s->w[21][c] = fu; s->w[22][c] = fv; s->w[23][c] = fe;
    }
}

static void kh_fill_prim_ghosts(ora_sim* s) {
    for (int y = -2; y < s->ny + 2; y++)
        for (int x = -2; x < s->nx + 2; x++) {
            if (x >= 0 && x < s->nx && y >= 0 && y < s->ny) continue;
            int mx, my;
            double sg;
            kh_map(s, x, y, &mx, &my, &sg);
            long c = pidx(s, x, y), m = pidx(s, mx, my);
            s->w[0][c] = s->w[0][m];
            s->w[1][c] = s->w[1][m];
            s->w[2][c] = sg * s->w[2][m];
            s->w[3][c] = s->w[3][m];
        }
}

/* periodic ghosts alias their real cell (stale or not); wall ghosts are
   computed from the mirrored primitives like any cell */
static void kh_fill_grad_ghosts(ora_sim* s) {
    for (int y = -1; y <= s->ny; y++)
        for (int x = -1; x <= s->nx; x++) {
            if (x >= 0 && x < s->nx && y >= 0 && y < s->ny) continue;
            int wall = s->p.bc_y == DBA_BC_WALL && (y < 0 || y >= s->ny);
            if (wall) {
                kh_grads_at(s, x, y);
            } else {
                int mx = map_i(x, s->nx, s->p.bc_x), my = map_i(y, s->ny, s->p.bc_y);
                long c = pidx(s, x, y), m = pidx(s, mx, my);
                for (int k = 4; k <= 15; k++) s->w[k][c] = s->w[k][m];
                s->w[28][c] = s->w[28][m];
            }
        }
}

static void kh_step(ora_sim* s) {
    const int dba = s->p.mode != DBA_MODE_REGULAR;
    /* 1. primitives on awake blocks (getPrimitive_wakeup) */
    for (int y = 0; y < s->ny; y++)
        for (int x = 0; x < s->nx; x++) {
            if (!active(s, 0, x, y, 0) && ORA_MUTANT != 3) continue;
            long i = (long)x + (long)y * s->nx, c = pidx(s, x, y);
            kh_prims_of(s, i, &s->w[0][c], &s->w[1][c], &s->w[2][c], &s->w[3][c]);
        }
    kh_fill_prim_ghosts(s);
    /* time step: fixed, or CFL over every cell's current primitives */
    if (s->p.dt_mode == DBA_DT_CFL) {
        double vmax = 0.0;
        for (int y = 0; y < s->ny; y++)
            for (int x = 0; x < s->nx; x++) {
                long c = pidx(s, x, y);
                double sp = kh_speed(s, s->w[0][c], s->w[1][c], s->w[2][c], s->w[3][c]);
                if (sp > vmax) vmax = sp;
            }
        s->dt_kh = s->p.cfl * (s->p.dx / vmax);
    } else {
        s->dt_kh = s->p.dt;
    }
    /* 2. gradients, primes, metric on awake blocks (extrap_in_time_wakeup) */
    for (int y = 0; y < s->ny; y++)
        for (int x = 0; x < s->nx; x++)
            if (active(s, 0, x, y, 0) || ORA_MUTANT == 3) kh_grads_at(s, x, y);
    kh_fill_grad_ghosts(s);
    /* 3. fluxes on awake blocks (calc_flux_wakeup); walls: the bottom face of row 0 */
    int wall = s->p.bc_y == DBA_BC_WALL;
    for (int y = 0; y < s->ny; y++)
        for (int x = 0; x < s->nx; x++) {
            if (!active(s, 0, x, y, 0) && ORA_MUTANT != 4) continue;
            kh_face_fluxes(s, x, y, 1, 1);
            if (wall && y == 0) kh_face_fluxes(s, x, -1, 0, 1);
        }
    /* 4. apply fluxes on awake blocks (apply_flux_wakeup) */
    double dtdx = s->dt_kh * s->p.dx;
    int cons = s->p.mode == DBA_MODE_CONSERVED;
    for (int y = 0; y < s->ny; y++)
        for (int x = 0; x < s->nx; x++) {
            long i = (long)x + (long)y * s->nx;
            if (!active(s, 0, x, y, 0)) {
                if (cons) for (int q = 0; q < 4; q++) s->w[24 + q][i] = 0.0;
                continue;
            }
            int xm = map_i(x - 1, s->nx, s->p.bc_x);
            long c = pidx(s, x, y), cxm = pidx(s, xm, y);
            long cym = (wall && y == 0) ? pidx(s, x, -1) : pidx(s, x, map_i(y - 1, s->ny, s->p.bc_y));
            /* U order Mass, Momx, Momy, Energy; flux arrays 16.. (x) and 20.. (y) in that order */
            for (int q = 0; q < 4; q++) {
                double ch = dtdx * (-1.0 * s->w[16 + q][c] + s->w[16 + q][cxm] - 1.0 * s->w[20 + q][c] + s->w[20 + q][cym]);
                if (cons) s->w[24 + q][i] = ch;
                s->f[q][i] += ch;
            }
        }
    if (!dba) return;
    if (cons) {
        /* example_Kelvin_Helmholtz/Conserved/KH_DBA.cu:1553-1575: every cell shifted */
        for (int q = 0; q < 4; q++) {
            double sum = 0.0;
            for (long i = 0; i < s->n; i++) sum += s->w[24 + q][i];
#if ORA_MUTANT == 5
            long na = 0;
            for (int y = 0; y < s->ny; y++)
                for (int x = 0; x < s->nx; x++) na += active(s, 0, x, y, 0);
            for (int y = 0; y < s->ny; y++)
                for (int x = 0; x < s->nx; x++)
                    if (active(s, 0, x, y, 0)) s->f[q][(long)x + (long)y * s->nx] -= sum / (double)na;
#else
            double avg = sum / (double)s->n;
            for (long i = 0; i < s->n; i++) s->f[q][i] -= avg;
#endif
        }
    }
    /* 5. Wakeup_Next: any(metric > Metric_eps), then 3x3 dilation */
    memset(s->next[0], 0, (size_t)s->nt);
    for (int y = 0; y < s->ny; y++)
        for (int x = 0; x < s->nx; x++) {
            int t = tile_of(s, x, y, 0);
            if (s->awake[0][t] && s->w[28][pidx(s, x, y)] > s->p.thresh_a) s->next[0][t] = 1;
        }
    dilate_set(s, 0);
}


/* ------------------------------------------------ reference 3D models (periodic) */

static inline long idx3(const ora_sim* s, int x, int y, int z) {
    return (long)map_i(x, s->nx, DBA_BC_PERIODIC) + (long)map_i(y, s->ny, DBA_BC_PERIODIC) * s->nx +
           (long)map_i(z, s->nz, DBA_BC_PERIODIC) * s->nx * s->ny;
}

/* Kobayashi-94 dendrite, example_Dendritic_growth/kernel.cu:181-293 (regular)
   and :322-634 (DBA). Pi = 3.14159 (kernel.cu:7); factors as kernel.cu:98-101.
   The powers pow(d,2), pow(d,4), pow(v2,2) are evaluated as products.
   Regular T: T + dt lapT + p_new - p (kernel.cu:289, left to right);
   DBA T: (T + dt lapT) + (p_new - p) (kernel.cu:471, 487).
   w[0] = metric_T. */
static void k94_step(ora_sim* s) {
    const double f0 = 1. / (2 * s->p.dx), f1 = 1. / s->p.dx / s->p.dx, f2 = -s->p.alpha / 3.14159, f3 = 1. / s->p.tau;
    const double dt = s->p.dt, eps = s->p.eps_bar, delta = s->p.delta, gam = s->p.gamma;
    const int dba = s->p.mode != DBA_MODE_REGULAR;
    const double* p = s->f[0];
    const double* T = s->f[1];
    double* pn = s->fn[0];
    double* Tn = s->fn[1];
    for (int z = 0; z < s->nz; z++)
        for (int y = 0; y < s->ny; y++)
            for (int x = 0; x < s->nx; x++) {
                long i = (long)x + (long)y * s->nx + (long)z * s->nx * s->ny;
                int ap = active(s, 0, x, y, z), at = active(s, 1, x, y, z);
                if (ap) {
                    double pxp = p[idx3(s, x + 1, y, z)], pxm = p[idx3(s, x - 1, y, z)];
                    double pyp = p[idx3(s, x, y + 1, z)], pym = p[idx3(s, x, y - 1, z)];
                    double pzp = p[idx3(s, x, y, z + 1)], pzm = p[idx3(s, x, y, z - 1)];
                    double dpx = (pxp - pxm) * f0, dpy = (pyp - pym) * f0, dpz = (pzp - pzm) * f0;
                    double lapP = (pxp + pxm + pyp + pym + pzp + pzm - 6 * p[i]) * f1;
                    double v2 = dpx * dpx + dpy * dpy + dpz * dpz;
                    double sigma;
                    if (v2 <= 1E-6)
                        sigma = 1;
                    else
                        sigma = 1 - delta * (1 - ((dpx * dpx) * (dpx * dpx) + (dpy * dpy) * (dpy * dpy) +
                                                  (dpz * dpz) * (dpz * dpz)) / (v2 * v2));
                    double m = f2 * atan(gam * sigma * T[i]);
                    double p_change = f3 * (eps * eps * lapP + p[i] * (1 - p[i]) * (p[i] - 0.5 + m));
                    pn[i] = p[i] + dt * p_change;
                }
                if (at || (!dba && ap)) {
                    double lapT = (T[idx3(s, x + 1, y, z)] + T[idx3(s, x - 1, y, z)] + T[idx3(s, x, y + 1, z)] +
                                   T[idx3(s, x, y - 1, z)] + T[idx3(s, x, y, z + 1)] + T[idx3(s, x, y, z - 1)] -
                                   6 * T[i]) * f1;
                    if (!dba) {
                        Tn[i] = T[i] + dt * lapT + pn[i] - p[i];
                    } else {
                        double tn = T[i] + dt * lapT;
                        if (ap) tn += pn[i] - p[i];
                        Tn[i] = tn;
                        s->w[0][i] = fabs(tn - T[i]);
                    }
                }
#if ORA_MUTANT == 2
                else if (dba && ap) {
                    Tn[i] = T[i] + (pn[i] - p[i]);
                }
#endif
            }
    if (!dba) {
        memcpy(s->f[0], s->fn[0], sizeof(double) * s->n);
        memcpy(s->f[1], s->fn[1], sizeof(double) * s->n);
        return;
    }
    memset(s->next[0], 0, (size_t)s->nt);
    memset(s->next[1], 0, (size_t)s->nt);
    double mp = s->p.thresh_a, mT = s->p.thresh_b;
    for (int z = 0; z < s->nz; z++)
        for (int y = 0; y < s->ny; y++)
            for (int x = 0; x < s->nx; x++) {
                long i = (long)x + (long)y * s->nx + (long)z * s->nx * s->ny;
                int t = tile_of(s, x, y, z);
                if (s->awake[0][t]) {
                    s->f[0][i] = pn[i];
                    if (pn[i] >= mp && pn[i] <= (1.0 - mp)) s->next[0][t] = 1;
                }
                if (s->awake[1][t]) {
                    s->f[1][i] = Tn[i];
                    if (s->w[0][i] >= mT) s->next[1][t] = 1;
                }
#if ORA_MUTANT == 2
                else if (s->awake[0][t]) {
                    s->f[1][i] = Tn[i];
                }
#endif
            }
    dilate_set(s, 0);
    dilate_set(s, 1);
}

/* 27-point isotropic Laplacian, example_Reactive_ModelB/kernel.cu:290-347, same summation order */
static double lap27(const ora_sim* s, const double* c, int x, int y, int z) {
#define C3(a, b, d) c[idx3(s, x + (a), y + (b), z + (d))]
    return 1.0 / (s->p.dx * s->p.dx) *
           (-64.0 / 15.0 * C3(0, 0, 0) +
            7.0 / 15.0 * (C3(1, 0, 0) + C3(-1, 0, 0) + C3(0, 1, 0) + C3(0, -1, 0) + C3(0, 0, 1) + C3(0, 0, -1)) +
            0.1 * (C3(1, 1, 0) + C3(1, -1, 0) + C3(-1, 1, 0) + C3(-1, -1, 0) + C3(1, 0, 1) + C3(0, 1, 1) +
                   C3(0, -1, 1) + C3(-1, 0, 1) + C3(1, 0, -1) + C3(0, 1, -1) + C3(0, -1, -1) + C3(-1, 0, -1)) +
            1.0 / 30.0 * (C3(1, 1, 1) + C3(-1, 1, 1) + C3(-1, -1, 1) + C3(1, -1, 1) + C3(1, 1, -1) + C3(-1, 1, -1) +
                          C3(-1, -1, -1) + C3(1, -1, -1)));
#undef C3
}

/* Ternary Flory-Huggins Model B with reaction, example_Reactive_ModelB/kernel.cu:349-437
   (regular), :469-763 (DBA), example_Reactive_ModelB/Conserved/kernel.cu:526-553, 630-702.
   f[0..3] = x, y, z, phi; w[0..2] = dfdx, dfdy, dfdz; w[3] = metric |lap z|;
   w[4..6] = change (conserved); w[7] = b. */
static void mb_step(ora_sim* s) {
    const dba_params* P = &s->p;
    const int dba = P->mode != DBA_MODE_REGULAR, cons = P->mode == DBA_MODE_CONSERVED;
    double *X = s->f[0], *Y = s->f[1], *Z = s->f[2], *PH = s->f[3];
    double *dfx = s->w[0], *dfy = s->w[1], *dfz = s->w[2], *B = s->w[7];
    for (int z = 0; z < s->nz; z++)
        for (int y = 0; y < s->ny; y++)
            for (int x = 0; x < s->nx; x++) {
                if (!active(s, 0, x, y, z) && ORA_MUTANT != 1) continue;
                long i = (long)x + (long)y * s->nx + (long)z * s->nx * s->ny;
                B[i] = 1.0 - X[i] - Y[i] - Z[i];
                dfx[i] = -1.0 + 1.0 / P->r_x + P->chi_xy * Y[i] + P->chi_xz * Z[i] + P->chi_xb * B[i] -
                         P->chi_xb * X[i] - P->chi_yb * Y[i] - P->chi_zb * Z[i] + log(X[i]) / P->r_x - log(B[i]) -
                         P->epsx_sq * lap27(s, X, x, y, z);
                dfy[i] = -1.0 + 1.0 / P->r_y + P->chi_xy * X[i] + P->chi_yz * Z[i] + P->chi_yb * B[i] -
                         P->chi_xb * X[i] - P->chi_yb * Y[i] - P->chi_zb * Z[i] + log(Y[i]) / P->r_y - log(B[i]) -
                         P->epsy_sq * lap27(s, Y, x, y, z);
                double lap_z = lap27(s, Z, x, y, z);
                dfz[i] = -1.0 + 1.0 / P->r_z + P->chi_xz * X[i] + P->chi_yz * Y[i] + P->chi_zb * B[i] -
                         P->chi_xb * X[i] - P->chi_yb * Y[i] - P->chi_zb * Z[i] + log(Z[i]) / P->r_z - log(B[i]) -
                         P->p_gel * PH[i] * PH[i] / 2.0 / (1.0 - P->z_crit) - P->epsz_sq * lap_z -
                         log(P->k_eq) / P->r_z;
                s->w[3][i] = fabs(lap_z);
            }
    for (int z = 0; z < s->nz; z++)
        for (int y = 0; y < s->ny; y++)
            for (int x = 0; x < s->nx; x++) {
                long i = (long)x + (long)y * s->nx + (long)z * s->nx * s->ny;
                if (!active(s, 0, x, y, z)) {
                    if (cons) s->w[4][i] = s->w[5][i] = s->w[6][i] = 0.0;
                    continue;
                }
                double R = P->k_0 * (exp(P->n_exp * P->v_x * dfx[i] + P->m_exp * P->v_y * dfy[i]) -
                                     exp(P->v_z * dfz[i]));
                double cx = P->dt * P->v_x * (P->mob_x * lap27(s, dfx, x, y, z) - P->n_exp * R);
                double cy = P->dt * P->v_y * (P->mob_y * lap27(s, dfy, x, y, z) - P->m_exp * R);
                double cz = P->dt * P->v_z * (P->mob_z * lap27(s, dfz, x, y, z) + R);
                s->fn[0][i] = X[i] + cx;
                s->fn[1][i] = Y[i] + cy;
                s->fn[2][i] = Z[i] + cz;
                if (cons) {
                    s->w[4][i] = cx;
                    s->w[5][i] = cy;
                    s->w[6][i] = cz;
                }
            }
    if (!dba) {
        for (int k = 0; k < 3; k++) memcpy(s->f[k], s->fn[k], sizeof(double) * s->n);
        return;
    }
    if (cons) {
#if ORA_MUTANT == 5
        long na = 0;
        for (int z = 0; z < s->nz; z++)
            for (int y = 0; y < s->ny; y++)
                for (int x = 0; x < s->nx; x++) na += active(s, 0, x, y, z);
        for (int k = 0; k < 3; k++) {
            double sum = 0.0;
            for (long i = 0; i < s->n; i++) sum += s->w[4 + k][i];
            for (int z = 0; z < s->nz; z++)
                for (int y = 0; y < s->ny; y++)
                    for (int x = 0; x < s->nx; x++)
                        if (active(s, 0, x, y, z)) {
                            long i = (long)x + (long)y * s->nx + (long)z * s->nx * s->ny;
                            s->f[k][i] = s->fn[k][i] - sum / (double)na;
                        }
        }
#else
        for (int k = 0; k < 3; k++) {
            double sum = 0.0;
            for (long i = 0; i < s->n; i++) sum += s->w[4 + k][i];
            double avg = sum / (double)s->n;
            for (long i = 0; i < s->n; i++) s->fn[k][i] -= avg;
        }
        for (int k = 0; k < 3; k++) memcpy(s->f[k], s->fn[k], sizeof(double) * s->n);
#endif
    } else {
        for (int z = 0; z < s->nz; z++)
            for (int y = 0; y < s->ny; y++)
                for (int x = 0; x < s->nx; x++)
                    if (active(s, 0, x, y, z)) {
                        long i = (long)x + (long)y * s->nx + (long)z * s->nx * s->ny;
                        for (int k = 0; k < 3; k++) s->f[k][i] = s->fn[k][i];
                    }
    }
    memset(s->next[0], 0, (size_t)s->nt);
    for (int z = 0; z < s->nz; z++)
        for (int y = 0; y < s->ny; y++)
            for (int x = 0; x < s->nx; x++) {
                int t = tile_of(s, x, y, z);
                if (s->awake[0][t] && s->w[3][(long)x + (long)y * s->nx + (long)z * s->nx * s->ny] > P->thresh_a)
                    s->next[0][t] = 1;
            }
    dilate_set(s, 0);
}

/* ------------------------------------------------------------------ public */

static void derive_constants(ora_sim* s) {
    const dba_params* p = &s->p;
    s->invdx2 = 1.0 / (p->dx * p->dx);
    s->inv6dx2 = 1.0 / (6.0 * p->dx * p->dx);
    s->inv2dx = 1.0 / (2.0 * p->dx);
    s->dt_tau = p->tau != 0.0 ? p->dt / p->tau : 0.0;
    s->alpha_pi = p->alpha / M_PI;
    double jf = (double)p->aniso_j;
    s->cos_jt0 = cos(jf * p->theta0);
    s->sin_jt0 = sin(jf * p->theta0);
    s->k_epsp = -(p->eps_bar * jf * p->delta);
}

static int nfields(int model) {
    switch (model) {
        case DBA_MODEL_CH: return 1;
        case DBA_MODEL_DENDRITE_K93: return 2;
        case DBA_MODEL_KH: return 4;
        case DBA_MODEL_DENDRITE_K94: return 2;
        case DBA_MODEL_MODELB: return 4;
        default: return 0;
    }
}

ora_sim* ora_create(const dba_params* p) {
    int nf = nfields(p->model);
    if (nf == 0) return NULL;
    if (p->nx <= 0 || p->ny <= 0 || p->nz <= 0) return NULL;
    if (p->nz != 1 && p->model != DBA_MODEL_DENDRITE_K94 && p->model != DBA_MODEL_MODELB) return NULL;
    ora_sim* s = (ora_sim*)calloc(1, sizeof(ora_sim));
    s->p = *p;
    s->nx = p->nx; s->ny = p->ny; s->nz = p->nz;
    s->n = (long)s->nx * s->ny * s->nz;
    s->tx = p->tile_x; s->ty = p->tile_y; s->tz = p->tile_z > 0 ? p->tile_z : 1;
    s->ntx = (s->nx + s->tx - 1) / s->tx;
    s->nty = (s->ny + s->ty - 1) / s->ty;
    s->ntz = (s->nz + s->tz - 1) / s->tz;
    s->nt = s->ntx * s->nty * s->ntz;
    derive_constants(s);
    for (int k = 0; k < nf; k++) {
        s->f[k] = (double*)calloc((size_t)s->n, sizeof(double));
        s->fn[k] = (double*)calloc((size_t)s->n, sizeof(double));
    }
    long nframe = (long)(s->nx + 4) * (s->ny + 4);
    for (int k = 0; k < MAXW; k++) s->w[k] = (double*)calloc((size_t)(nframe > s->n ? nframe : s->n), sizeof(double));
    for (int k = 0; k < 2; k++) {
        s->awake[k] = (uint8_t*)malloc((size_t)s->nt);
        memset(s->awake[k], 1, (size_t)s->nt);  /* all blocks awake at start */
        s->next[k] = (uint8_t*)calloc((size_t)s->nt, 1);
    }
    s->hist_cap = p->history_cap > 0 ? p->history_cap : 0;
    s->hist = (int32_t*)calloc((size_t)(2 * s->hist_cap + 2), sizeof(int32_t));
    return s;
}

void ora_destroy(ora_sim* s) {
    if (!s) return;
    for (int k = 0; k < MAXF; k++) { free(s->f[k]); free(s->fn[k]); }
    for (int k = 0; k < MAXW; k++) free(s->w[k]);
    for (int k = 0; k < 2; k++) { free(s->awake[k]); free(s->next[k]); }
    free(s->hist);
    free(s);
}

int ora_set_field(ora_sim* s, int field, const double* src) {
    if (field < 0 || field >= nfields(s->p.model)) return -1;
    memcpy(s->f[field], src, sizeof(double) * s->n);
    memcpy(s->fn[field], src, sizeof(double) * s->n);
    return 0;
}

int ora_get_field(ora_sim* s, int field, double* dst) {
    if (field < 0 || field >= nfields(s->p.model)) return -1;
    memcpy(dst, s->f[field], sizeof(double) * s->n);
    return 0;
}

int ora_step(ora_sim* s, int nsteps) {
    for (int k = 0; k < nsteps; k++) {
        record_history(s);
        switch (s->p.model) {
            case DBA_MODEL_CH: ch_step(s); break;
            case DBA_MODEL_DENDRITE_K93: k93_step(s); break;
            case DBA_MODEL_KH: kh_step(s); break;
            case DBA_MODEL_DENDRITE_K94: k94_step(s); break;
            case DBA_MODEL_MODELB: mb_step(s); break;
            default: return -1;
        }
        s->steps++;
        s->last_dt = s->p.model == DBA_MODEL_KH ? s->dt_kh : s->p.dt;
        s->time += s->last_dt;
    }
    return 0;
}

int ora_get_awake(ora_sim* s, int set, uint8_t* dst) {
    if (set < 0 || set > 1) return -1;
    if (s->p.mode == DBA_MODE_REGULAR) memset(dst, 1, (size_t)s->nt);
    else memcpy(dst, s->awake[set], (size_t)s->nt);
    return 0;
}

int ora_set_awake(ora_sim* s, int set, const uint8_t* src) {
    if (set < 0 || set > 1) return -1;
    memcpy(s->awake[set], src, (size_t)s->nt);
    return 0;
}

int ora_get_next(ora_sim* s, int set, uint8_t* dst) {
    if (set < 0 || set > 1) return -1;
    memcpy(dst, s->next[set], (size_t)s->nt);
    return 0;
}

int ora_ntiles(ora_sim* s) { return s->nt; }
double ora_time(ora_sim* s) { return s->time; }
double ora_last_dt(ora_sim* s) { return s->last_dt; }

int ora_history(ora_sim* s, int32_t* out, int cap) {
    int n = s->hist_n < cap ? s->hist_n : cap;
    memcpy(out, s->hist, sizeof(int32_t) * 2 * (size_t)n);
    return n;
}

/* -------------------------------------------- threaded regular-mode baseline */
/* y-slab decomposition over pthreads with a barrier between phases, the
   shared-memory analogue of the reference MPI_version z/y slabs
   (example_Reactive_ModelB/CPU_version/MPI_version/base_CPU.cpp:182-226). */

typedef struct {
    ora_sim* s;
    int y0, y1, nsteps;
    int tid, nth;
    pthread_barrier_t* bar;
} slab_arg;

static void slab_copy(ora_sim* s, int nf, int y0, int y1) {
    for (int k = 0; k < nf; k++)
        memcpy(s->f[k] + (long)y0 * s->nx, s->fn[k] + (long)y0 * s->nx, sizeof(double) * (size_t)(y1 - y0) * s->nx);
}

/* KH regular step over a y-slab of a thread team: the phases of kh_step with
   barriers between them; the ghost fills and the CFL reduction are done by
   thread 0 (y0 == 0). Every cell sees the same operations as in kh_step. */
static double* kh_slab_vmax;  /* per-thread maxima, indexed by thread */

static void kh_step_slab(ora_sim* s, int y0, int y1, int tid, int nth, pthread_barrier_t* bar) {
    for (int y = y0; y < y1; y++)
        for (int x = 0; x < s->nx; x++) {
            long i = (long)x + (long)y * s->nx, c = pidx(s, x, y);
            kh_prims_of(s, i, &s->w[0][c], &s->w[1][c], &s->w[2][c], &s->w[3][c]);
        }
    if (s->p.dt_mode == DBA_DT_CFL) {
        double vmax = 0.0;
        for (int y = y0; y < y1; y++)
            for (int x = 0; x < s->nx; x++) {
                long c = pidx(s, x, y);
                double sp = kh_speed(s, s->w[0][c], s->w[1][c], s->w[2][c], s->w[3][c]);
                if (sp > vmax) vmax = sp;
            }
        kh_slab_vmax[tid] = vmax;
    }
    pthread_barrier_wait(bar);
    if (tid == 0) {
        kh_fill_prim_ghosts(s);
        if (s->p.dt_mode == DBA_DT_CFL) {
            double vmax = 0.0;
            for (int k = 0; k < nth; k++)
                if (kh_slab_vmax[k] > vmax) vmax = kh_slab_vmax[k];
            s->dt_kh = s->p.cfl * (s->p.dx / vmax);
        } else {
            s->dt_kh = s->p.dt;
        }
    }
    pthread_barrier_wait(bar);
    for (int y = y0; y < y1; y++)
        for (int x = 0; x < s->nx; x++) kh_grads_at(s, x, y);
    pthread_barrier_wait(bar);
    if (tid == 0) kh_fill_grad_ghosts(s);
    pthread_barrier_wait(bar);
    int wall = s->p.bc_y == DBA_BC_WALL;
    for (int y = y0; y < y1; y++)
        for (int x = 0; x < s->nx; x++) {
            kh_face_fluxes(s, x, y, 1, 1);
            if (wall && y == 0) kh_face_fluxes(s, x, -1, 0, 1);
        }
    pthread_barrier_wait(bar);
    double dtdx = s->dt_kh * s->p.dx;
    for (int y = y0; y < y1; y++)
        for (int x = 0; x < s->nx; x++) {
            long i = (long)x + (long)y * s->nx;
            int xm = map_i(x - 1, s->nx, s->p.bc_x);
            long c = pidx(s, x, y), cxm = pidx(s, xm, y);
            long cym = (wall && y == 0) ? pidx(s, x, -1) : pidx(s, x, map_i(y - 1, s->ny, s->p.bc_y));
            for (int q = 0; q < 4; q++)
                s->f[q][i] += dtdx * (-1.0 * s->w[16 + q][c] + s->w[16 + q][cxm] - 1.0 * s->w[20 + q][c] + s->w[20 + q][cym]);
        }
}

static void* slab_worker(void* argp) {
    slab_arg* a = (slab_arg*)argp;
    ora_sim* s = a->s;
    if (s->p.model == DBA_MODEL_KH) {
        for (int t = 0; t < a->nsteps; t++) {
            kh_step_slab(s, a->y0, a->y1, a->tid, a->nth, a->bar);
            pthread_barrier_wait(a->bar);
        }
        return NULL;
    }
    for (int t = 0; t < a->nsteps; t++) {
        if (s->p.model == DBA_MODEL_CH) {
            ch_calc_mu(s, a->y0, a->y1);
            pthread_barrier_wait(a->bar);
            ch_update(s, a->y0, a->y1);
            pthread_barrier_wait(a->bar);
            slab_copy(s, 1, a->y0, a->y1);
        } else {
            k93_finite_diff(s, a->y0, a->y1);
            pthread_barrier_wait(a->bar);
            k93_evolve(s, a->y0, a->y1);
            pthread_barrier_wait(a->bar);
            slab_copy(s, 2, a->y0, a->y1);
        }
        pthread_barrier_wait(a->bar);
    }
    return NULL;
}

double ora_bench(ora_sim* s, int nsteps, int nthreads) {
    if (s->p.mode != DBA_MODE_REGULAR) return -1.0;
    if (s->p.model == DBA_MODEL_KH && nthreads <= 1) {
        /* serial: the oracle's own step */
        struct timespec a, b;
        clock_gettime(CLOCK_MONOTONIC, &a);
        for (int k = 0; k < nsteps; k++) kh_step(s);
        clock_gettime(CLOCK_MONOTONIC, &b);
        s->steps += nsteps;
        return (double)(b.tv_sec - a.tv_sec) + 1e-9 * (double)(b.tv_nsec - a.tv_nsec);
    }
    if (s->p.model == DBA_MODEL_DENDRITE_K94 || s->p.model == DBA_MODEL_MODELB) {
        /* 3D reference models: serial step loop */
        struct timespec a, b;
        clock_gettime(CLOCK_MONOTONIC, &a);
        ora_step(s, nsteps);
        clock_gettime(CLOCK_MONOTONIC, &b);
        return (double)(b.tv_sec - a.tv_sec) + 1e-9 * (double)(b.tv_nsec - a.tv_nsec);
    }
    if (s->p.model != DBA_MODEL_KH && s->p.model != DBA_MODEL_CH && s->p.model != DBA_MODEL_DENDRITE_K93)
        return -1.0;
    if (nthreads < 1) nthreads = 1;
    if (nthreads > s->ny) nthreads = s->ny;
    if (nthreads > 256) nthreads = 256;
    struct timespec t0, t1;
    pthread_barrier_t bar;
    pthread_barrier_init(&bar, NULL, (unsigned)nthreads);
    pthread_t* th = (pthread_t*)malloc(sizeof(pthread_t) * (size_t)nthreads);
    slab_arg* args = (slab_arg*)malloc(sizeof(slab_arg) * (size_t)nthreads);
    for (int k = 0; k < nthreads; k++) {
        args[k].s = s;
        args[k].y0 = (int)((long)s->ny * k / nthreads);
        args[k].y1 = (int)((long)s->ny * (k + 1) / nthreads);
        args[k].nsteps = nsteps;
        args[k].bar = &bar;
        args[k].tid = k;
        args[k].nth = nthreads;
    }
    double vm[256];
    kh_slab_vmax = vm;
    clock_gettime(CLOCK_MONOTONIC, &t0);
    for (int k = 1; k < nthreads; k++) pthread_create(&th[k], NULL, slab_worker, &args[k]);
    slab_worker(&args[0]);
    for (int k = 1; k < nthreads; k++) pthread_join(th[k], NULL);
    clock_gettime(CLOCK_MONOTONIC, &t1);
    pthread_barrier_destroy(&bar);
    free(th);
    free(args);
    s->steps += nsteps;
    return (double)(t1.tv_sec - t0.tv_sec) + 1e-9 * (double)(t1.tv_nsec - t0.tv_nsec);
}
