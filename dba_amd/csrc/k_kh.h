// k_kh.h - fused DBA step for 2D compressible Euler (Kelvin-Helmholtz):
// MUSCL-Hancock with central gradients, half-step predictor and Rusanov fluxes.
//
// Reference counterpart: getPrimitive_wakeup -> extrap_in_time_wakeup ->
// calc_flux_wakeup -> apply_flux_wakeup -> Wakeup_Next
// (example_Kelvin_Helmholtz/KH_DBA.cu:418-985), conserved variant
// example_Kelvin_Helmholtz/Conserved/KH_DBA.cu:832-955. Same arithmetic
// (getFlux :662-697, extrapolateInSpaceToFace :630-660, applyFluxes :818-841),
// generalised to a gas gamma, a CFL time step and reflective y walls (mirrored
// ghost cells with the normal velocity negated).
//
// The reference reads the primitives, gradients, primes and fluxes that
// sleeping neighbours left behind. Only a tile's edges are ever read that way,
// so each awake tile stores exactly those: primitives on its ring, (primes,
// x-gradients) on its left column, (primes, y-gradients) on its bottom row, the
// x-face flux of its right column and the y-face flux of its top row.
// Everything else is recomputed in LDS from the conserved state of the tile
// and its neighbours (2-cell halo). The stored state is one contiguous record
// per tile (KhEdge), so a tile's edge writes coalesce instead of landing one
// 8-byte value per cache line down its columns.
//
// LDS (TB = 16): primitives 4 x 20^2, primes / x- / y-gradients 12 x 18^2;
// the x-face fluxes reuse the primitive space, the y-face fluxes the
// x-gradient space.
#pragma once
#include "common.h"

namespace dba {

struct KhC {
    double g, dx, dy, dt, vol;
    double idx, idy, ivol;  // 1/dx, 1/dy, 1/vol: exact when P2 (below)
};

// x / d. P2: d is a power of two (the grid spacings and the cell area of a
// power-of-two grid on a unit square), so x * (1/d) rounds the same exact real
// as x / d and gives the same double in every case, at the cost of a multiply
// instead of a correctly rounded division.
template <bool P2>
__device__ __forceinline__ double kdiv(double x, double d, double id) {
    if constexpr (P2) return x * id;
    else return x / d;
}

template <bool P2 = false>
__device__ __forceinline__ void kh_prims(const KhC& c, double m, double mx, double my, double e, double* r,
                                         double* u, double* v, double* p) {
    const double rho = kdiv<P2>(m, c.vol, c.ivol);
    const double uu = kdiv<P2>(mx / rho, c.vol, c.ivol);
    const double vv = kdiv<P2>(my / rho, c.vol, c.ivol);
    *r = rho;
    *u = uu;
    *v = vv;
    *p = (kdiv<P2>(e, c.vol, c.ivol) - 0.5 * rho * (uu * uu + vv * vv)) * (c.g - 1.0);
}

__device__ __forceinline__ double kh_speed(const KhC& c, double r, double u, double v, double p) {
    return sqrt(c.g * p / r) + sqrt(u * u + v * v);
}

// getFlux (example_Kelvin_Helmholtz/KH_DBA.cu:662-697)
__device__ __forceinline__ void kh_rusanov(const KhC& c, double rL, double rR, double pL, double pR, double uL,
                                           double uR, double vL, double vR, double* fm, double* fe, double* fu,
                                           double* fv) {
    const double g = c.g;
    const double en_L = pL / (g - 1.0) + 0.5 * rL * (uL * uL + vL * vL);
    const double en_R = pR / (g - 1.0) + 0.5 * rR * (uR * uR + vR * vR);
    const double rho_star = 0.5 * (rL + rR);
    const double momx_star = 0.5 * (rL * uL + rR * uR);
    const double momy_star = 0.5 * (rL * vL + rR * vR);
    const double en_star = 0.5 * (en_L + en_R);
    const double p_star = (g - 1.0) * (en_star - 0.5 * (momx_star * momx_star + momy_star * momy_star) / rho_star);
    const double flux_Mass = momx_star;
    const double flux_Momx = momx_star * momx_star / rho_star + p_star;
    const double flux_Momy = momx_star * momy_star / rho_star;
    const double flux_Energy = (en_star + p_star) * momx_star / rho_star;
    const double C_L = sqrt(g * pL / rL) + fabs(uL);
    const double C_R = sqrt(g * pR / rR) + fabs(uR);
    const double C = (C_L >= C_R) ? C_L : C_R;
    *fm = flux_Mass - C * 0.5 * (rL - rR);
    *fu = flux_Momx - C * 0.5 * (rL * uL - rR * uR);
    *fv = flux_Momy - C * 0.5 * (rL * vL - rR * vR);
    *fe = flux_Energy - C * 0.5 * (en_L - en_R);
}

// Edge record of a tile, FP64, field-major: ring primitives r, u, v, p (ring
// cell order: bottom row, top row, left column, right column, the columns
// without their corners); left column primes + x-gradients; bottom row primes +
// y-gradients; x-face fluxes (m, u, v, e) of the right column; y-face fluxes of
// the top row.
template <int TB>
struct KhEdge {
    static constexpr int RING = 4 * TB - 4;
    static constexpr int LEFT = 4 * RING;       // [8][TB]: primes, x-gradients by row
    static constexpr int BOTTOM = LEFT + 8 * TB;  // [8][TB]: primes, y-gradients by column
    static constexpr int XF = BOTTOM + 8 * TB;  // [4][TB] by row
    static constexpr int YF = XF + 4 * TB;      // [4][TB] by column
    static constexpr int R = YF + 4 * TB;       // doubles per tile: 40 TB - 16
    // slot of tile-local cell (lx, ly) on the ring, -1 inside
    __device__ __forceinline__ static int ring(int lx, int ly) {
        if (ly == 0) return lx;
        if (ly == TB - 1) return TB + lx;
        if (lx == 0) return 2 * TB + ly - 1;
        if (lx == TB - 1) return 3 * TB - 2 + ly - 1;
        return -1;
    }
    // the cell of ring slot k (the inverse of ring)
    __device__ __forceinline__ static void cell(int k, int* lx, int* ly) {
        if (k < TB) { *lx = k; *ly = 0; }
        else if (k < 2 * TB) { *lx = k - TB; *ly = TB - 1; }
        else if (k < 3 * TB - 2) { *lx = 0; *ly = k - 2 * TB + 1; }
        else { *lx = TB - 1; *ly = k - (3 * TB - 2) + 1; }
    }
};

template <int TB, int NT, bool P2>
__global__ __launch_bounds__(NT) void kh_step_kernel(const WorkList wl) {
    using E = KhEdge<TB>;
    constexpr int W2 = TB + 4;
    constexpr int W1 = TB + 2;
    constexpr int A2 = W2 * W2;
    constexpr int A1 = W1 * W1;
    constexpr int FW = TB + 1;       // flux array row width (x faces) / rows (y faces)
    constexpr int NF = FW * TB;      // faces per direction
    constexpr int YS = (NF + NT - 1) / NT;
    static_assert(A2 >= NF, "y-face fluxes reuse the primitive space");
    // The gradients are not stored: the flux phases recompute them from the
    // primitives with the expression phase 2 uses (the same doubles), except the
    // stale ones a sleeping right / top neighbour kept. 33 KB at TB = 16 (4
    // workgroups per CU instead of 3 with stored gradients).
    __shared__ double sP[4 * A2];    // primitives r, vx, vy, p; later y-face fluxes
    __shared__ double sPR[4 * A1];   // primes
    __shared__ double sXF[4 * NF];   // x-face fluxes
    __shared__ double sGXR[4][TB];   // x-gradients of a sleeping right neighbour's left column
    __shared__ double sGYT[4][TB];   // y-gradients of a sleeping top neighbour's bottom row
    __shared__ double sred[NT / 64 + 1];
    __shared__ int sflag[NT / 64 + 1];
    __shared__ NbTable nb;

    WorkCursor cur;
    work_begin(wl, cur, true);
    const int total = wl.offsets[wl.nseg];
    const int tid = threadIdx.x;
    for (;;) {
        CInst* Ip;
        int t;
        uint64_t rec;
        if (!work_next(wl, cur, total, &Ip, &t, &rec)) break;
        CInst& I = *Ip;
        const int nx = I.nx, ny = I.ny, ntx = I.ntx;
        const int x0 = (t % ntx) * TB, y0 = (t / ntx) * TB;
        const bool regular = I.mode == 0;
        const bool dba = !regular;
        const int ncons = I.ncons;
        const bool wall = I.bc_y == 2;
        const bool cfl = I.dt_mode == 1;
        KhC c;
        c.g = I.gas_gamma;
        c.dx = I.dx;
        c.dy = I.dy;
        c.vol = I.vol;
        c.idx = 1.0 / c.dx;
        c.idy = 1.0 / c.dy;
        c.ivol = 1.0 / c.vol;
        c.dt = cfl ? I.scal[1] : I.dt;
        nb_fill<TB>(nb, I, t, x0, y0, regular, 1, ncons, rec);
        __syncthreads();
        const int pself = nb.par[0][4];

        // 1. primitives on the tile + 2-cell halo: from U where the owner runs
        //    this step, else the primitives its ring kept (a sleeping owner's
        //    inner cells are never reached by a stencil of a running cell). Every
        //    load of the thread's cells is issued before the first wait: the
        //    source address is selected per cell (U of the owner's buffer, or its
        //    ring record), a cell past the region loads cell 0 again, and a cell
        //    with no ring value loads the record's first word; neither is stored.
        {
            constexpr int K1 = (A2 + NT - 1) / NT;
            const gdouble *u00 = I.buf[0][0], *u01 = I.buf[0][1], *u02 = I.buf[0][2], *u03 = I.buf[0][3];
            const gdouble *u10 = I.buf[1][0], *u11 = I.buf[1][1], *u12 = I.buf[1][2], *u13 = I.buf[1][3];
            const gdouble* rec0 = I.aux[0];
            asm volatile("" : "+s"(u00), "+s"(u01), "+s"(u02), "+s"(u03), "+s"(u10), "+s"(u11), "+s"(u12),
                         "+s"(u13), "+s"(rec0));
            double val[K1][4];
            int dq[K1], riq[K1];
            double sgq[K1];
#pragma unroll
            for (int q = 0; q < K1; q++) {
                const int k0 = tid + q * NT, k = k0 < A2 ? k0 : 0;
                const int lx = k % W2, ly = k / W2;
                const int d = nb_dir1<TB, 2>(lx) + 3 * nb_dir1<TB, 2>(ly);
                const int gy = y0 + ly - 2;
                const int mx = map_index(x0 + lx - 2, nx, 0), my = map_index(gy, ny, wall ? 1 : 0);
                const int64_t gi = mx + (int64_t)my * nx;
                const int ri = E::ring(mx & (TB - 1), my & (TB - 1));
                const bool act = nb.act[0][d] != 0;
                const bool pp = nb.par[0][d] != 0;
                const gdouble* e = rec0 + (int64_t)nb.tile[d] * E::R + (ri >= 0 ? ri : 0);
                val[q][0] = act ? (pp ? u10 : u00)[gi] : e[0];
                val[q][1] = act ? (pp ? u11 : u01)[gi] : e[E::RING];
                val[q][2] = act ? (pp ? u12 : u02)[gi] : e[2 * E::RING];
                val[q][3] = act ? (pp ? u13 : u03)[gi] : e[3 * E::RING];
                dq[q] = d;
                riq[q] = ri;
                sgq[q] = (wall && (gy < 0 || gy >= ny)) ? -1.0 : 1.0;
            }
#pragma unroll
            for (int q = 0; q < K1; q++) {
                const int k = tid + q * NT;
                if (k >= A2) continue;
                const int d = dq[q];
                double r, u, v, p;
                if (nb.act[0][d]) {
                    const double um = val[q][0] - nb.shift[0][d];
                    const double ux = val[q][1] - nb.shift[1][d];
                    const double uy = val[q][2] - nb.shift[2][d];
                    const double ue = val[q][3] - nb.shift[3][d];
                    kh_prims<P2>(c, um, ux, uy, ue, &r, &u, &v, &p);
                } else if (riq[q] >= 0) {
                    r = val[q][0];
                    u = val[q][1];
                    v = val[q][2];
                    p = val[q][3];
                } else {
                    r = u = v = p = 0.0;
                }
                sP[0 * A2 + k] = r;
                sP[1 * A2 + k] = u;
                sP[2 * A2 + k] = sgq[q] * v;
                sP[3 * A2 + k] = p;
            }
        }
        __syncthreads();

        // stale speed (primitives at the start of the step)
        double vstale = 0.0;
        if (cfl) {
            for (int k = tid; k < TB * TB; k += NT) {
                const int lx = k % TB, ly = k / TB;
                const int q = (lx + 2) + (ly + 2) * W2;
                vstale = fmax(vstale, kh_speed(c, sP[q], sP[A2 + q], sP[2 * A2 + q], sP[3 * A2 + q]));
            }
        }

        // 2. gradients, primes on the tile + ring (no ring corners); metric on the tile
        int flag = 0;
        for (int k = tid; k < A1; k += NT) {
            const int lx = k % W1, ly = k / W1;
            const bool cx = lx == 0 || lx == TB + 1, cy = ly == 0 || ly == TB + 1;
            if (cx && cy) continue;
            const int d = nb_dir1<TB, 1>(lx) + 3 * nb_dir1<TB, 1>(ly);
            if (!nb.act[0][d]) {
                // sleeping neighbour: its stored edge state (right: primes + x-grads
                // of its left column; top: primes + y-grads of its bottom row)
                if (lx == TB + 1 || ly == TB + 1) {
                    const gdouble* e = I.aux[0] + (int64_t)nb.tile[d] * E::R;
                    if (lx == TB + 1) {
                        const int row = map_index(y0 + ly - 1, ny, wall ? 1 : 0) & (TB - 1);
                        for (int q = 0; q < 4; q++) {
                            sPR[q * A1 + k] = e[E::LEFT + q * TB + row];
                            sGXR[q][ly - 1] = e[E::LEFT + (4 + q) * TB + row];
                        }
                    } else {
                        const int col = map_index(x0 + lx - 1, nx, 0) & (TB - 1);
                        for (int q = 0; q < 4; q++) {
                            sPR[q * A1 + k] = e[E::BOTTOM + q * TB + col];
                            sGYT[q][lx - 1] = e[E::BOTTOM + (4 + q) * TB + col];
                        }
                    }
                }
                continue;
            }
            const int p0 = (lx + 1) + (ly + 1) * W2;
            const int pxm = p0 - 1, pxp = p0 + 1, pym = p0 - W2, pyp = p0 + W2;
            // Grad_x / Grad_y (example_Kelvin_Helmholtz/KH_DBA.cu:249-275)
            const double r_x = kdiv<P2>((sP[pxp] - sP[pxm]) / 2.0, c.dx, c.idx), r_y = kdiv<P2>((sP[pyp] - sP[pym]) / 2.0, c.dy, c.idy);
            const double u_x = kdiv<P2>((sP[A2 + pxp] - sP[A2 + pxm]) / 2.0, c.dx, c.idx);
            const double u_y = kdiv<P2>((sP[A2 + pyp] - sP[A2 + pym]) / 2.0, c.dy, c.idy);
            const double v_x = kdiv<P2>((sP[2 * A2 + pxp] - sP[2 * A2 + pxm]) / 2.0, c.dx, c.idx);
            const double v_y = kdiv<P2>((sP[2 * A2 + pyp] - sP[2 * A2 + pym]) / 2.0, c.dy, c.idy);
            const double p_x = kdiv<P2>((sP[3 * A2 + pxp] - sP[3 * A2 + pxm]) / 2.0, c.dx, c.idx);
            const double p_y = kdiv<P2>((sP[3 * A2 + pyp] - sP[3 * A2 + pym]) / 2.0, c.dy, c.idy);
            const double r = sP[p0], u = sP[A2 + p0], v = sP[2 * A2 + p0], p = sP[3 * A2 + p0];
            const double dt = c.dt, g = c.g;
            // extrap_in_time (example_Kelvin_Helmholtz/KH_DBA.cu:524-531)
            const double rp = r - 0.5 * dt * (u * r_x + r * u_x + v * r_y + r * v_y);
            const double up = u - 0.5 * dt * (u * u_x + v * u_y + (1.0 / r) * p_x);
            const double vp = v - 0.5 * dt * (u * v_x + v * v_y + (1.0 / r) * p_y);
            const double pp = p - 0.5 * dt * (g * p * (u_x + v_y) + u * p_x + v * p_y);
            sPR[k] = rp;
            sPR[A1 + k] = up;
            sPR[2 * A1 + k] = vp;
            sPR[3 * A1 + k] = pp;
            if (d == 4) {
                const double metric = I.metric_kind == 1
                    ? fabs(r_x) + fabs(r_y)
                    : fabs(r_x) + fabs(r_y) + fabs(p_x) + fabs(p_y) + fabs(u_x) + fabs(u_y) + fabs(v_x) + fabs(v_y);
                flag |= metric > I.thresh_a;
            }
        }
        __syncthreads();

        // Grad_x / Grad_y of W1 cell (lx, ly) recomputed from the primitives
        // (example_Kelvin_Helmholtz/KH_DBA.cu:249-275, the expressions of phase 2)
        auto grad_x = [&](int q, int lx, int ly) {
            const int p0 = (lx + 1) + (ly + 1) * W2;
            return kdiv<P2>((sP[q * A2 + p0 + 1] - sP[q * A2 + p0 - 1]) / 2.0, c.dx, c.idx);
        };
        auto grad_y = [&](int q, int lx, int ly) {
            const int p0 = (lx + 1) + (ly + 1) * W2;
            return kdiv<P2>((sP[q * A2 + p0 + W2] - sP[q * A2 + p0 - W2]) / 2.0, c.dy, c.idy);
        };

        // detector; the edge record for sleeping-neighbour reads only when the
        // tile did not fire: a tile that fired is awake next step (the dilation
        // keeps the tile itself) and a record is read only while its owner
        // sleeps, so the record a sleeping owner presents is its last step's
        const int any = dba ? block_any<NT>(flag, sflag) : 0;
        const bool store = dba && !any;
        if (store) {
            gdouble* e = I.aux[0] + (int64_t)t * E::R;
            // ring primitives
            for (int k = tid; k < E::RING; k += NT) {
                int lx, ly;
                E::cell(k, &lx, &ly);
                const int q = (lx + 2) + (ly + 2) * W2;
                e[k] = sP[q];
                e[E::RING + k] = sP[A2 + q];
                e[2 * E::RING + k] = sP[2 * A2 + q];
                e[3 * E::RING + k] = sP[3 * A2 + q];
            }
            // left column: primes + x-gradients; bottom row: primes + y-gradients
            for (int k = tid; k < 2 * 8 * TB; k += NT) {
                const int part = k / (8 * TB), j = k % (8 * TB), q = j / TB, i = j % TB;
                const int lx = part == 0 ? 1 : 1 + i, ly = part == 0 ? 1 + i : 1;  // W1 coordinates
                const double v = q < 4 ? sPR[q * A1 + lx + ly * W1]
                                       : (part == 0 ? grad_x(q - 4, lx, ly) : grad_y(q - 4, lx, ly));
                e[(part == 0 ? E::LEFT : E::BOTTOM) + j] = v;
            }
        }

        // 3a. x-face fluxes of W1 columns 0..TB on the tile rows
        for (int k = tid; k < NF; k += NT) {
            const int lx = k % FW, ly = 1 + k / FW;  // W1 coordinates of the cell left of the face
            const int a = lx + ly * W1, b = a + 1;
            double fm, fe, fu, fv;
            const int d = nb_dir1<TB, 1>(lx) + 3 * nb_dir1<TB, 1>(ly);
            if (lx == 0 && !nb.act[0][d]) {
                // the x-face fluxes the sleeping left neighbour kept for its right column
                const gdouble* e = I.aux[0] + (int64_t)nb.tile[d] * E::R + E::XF + (ly - 1);
                fm = e[0];
                fu = e[TB];
                fv = e[2 * TB];
                fe = e[3 * TB];
            } else {
                // the right cell is a sleeping right neighbour's: its stored gradients
                const bool bstale = lx == TB && !nb.act[0][nb_dir1<TB, 1>(TB + 1) + 3 * nb_dir1<TB, 1>(ly)];
                double L[4], R[4];
                for (int q = 0; q < 4; q++) {
                    const double gb = bstale ? sGXR[q][ly - 1] : grad_x(q, lx + 1, ly);
                    L[q] = sPR[q * A1 + b] - gb * c.dx / 2.0;                    // face[0] of the right cell
                    R[q] = sPR[q * A1 + a] + grad_x(q, lx, ly) * c.dx / 2.0;  // face[1] of the left cell
                }
                kh_rusanov(c, L[0], R[0], L[3], R[3], L[1], R[1], L[2], R[2], &fm, &fe, &fu, &fv);
                if (store && lx == TB) {
                    gdouble* e = I.aux[0] + (int64_t)t * E::R + E::XF + (ly - 1);
                    e[0] = fm;
                    e[TB] = fu;
                    e[2 * TB] = fv;
                    e[3 * TB] = fe;
                }
            }
            sXF[0 * NF + k] = fm;
            sXF[1 * NF + k] = fu;
            sXF[2 * NF + k] = fv;
            sXF[3 * NF + k] = fe;
        }
        // 3b. y-face fluxes of W1 rows 0..TB on the tile columns, held in registers
        //     until every thread has read the primitives, then stored over them
        double yf[YS][4];
#pragma unroll
        for (int i = 0; i < YS; i++) {
            const int k = tid + i * NT;
            if (k >= NF) break;
            const int ly = k / TB, lx = 1 + k % TB;  // W1 coordinates of the cell below the face
            const int a = lx + ly * W1, b = a + W1;
            double fm, fe, fu, fv;
            const int d = nb_dir1<TB, 1>(lx) + 3 * nb_dir1<TB, 1>(ly);
            if (ly == 0 && !nb.act[0][d]) {
                // the y-face fluxes the sleeping lower neighbour kept for its top row
                const gdouble* e = I.aux[0] + (int64_t)nb.tile[d] * E::R + E::YF + (lx - 1);
                fm = e[0];
                fu = e[TB];
                fv = e[2 * TB];
                fe = e[3 * TB];
            } else {
                const bool bstale = ly == TB && !nb.act[0][nb_dir1<TB, 1>(lx) + 3 * nb_dir1<TB, 1>(TB + 1)];
                double L[4], R[4];
                for (int q = 0; q < 4; q++) {
                    const double gb = bstale ? sGYT[q][lx - 1] : grad_y(q, lx, ly + 1);
                    L[q] = sPR[q * A1 + b] - gb * c.dy / 2.0;                    // face[2] of the upper cell
                    R[q] = sPR[q * A1 + a] + grad_y(q, lx, ly) * c.dy / 2.0;  // face[3] of the lower cell
                }
                // momentum roles swapped for y faces (example_Kelvin_Helmholtz/KH_DBA.cu:738-740)
                kh_rusanov(c, L[0], R[0], L[3], R[3], L[2], R[2], L[1], R[1], &fm, &fe, &fv, &fu);
                if (store && ly == TB) {
                    gdouble* e = I.aux[0] + (int64_t)t * E::R + E::YF + (lx - 1);
                    e[0] = fm;
                    e[TB] = fu;
                    e[2 * TB] = fv;
                    e[3 * TB] = fe;
                }
            }
            yf[i][0] = fm;
            yf[i][1] = fu;
            yf[i][2] = fv;
            yf[i][3] = fe;
        }
        __syncthreads();
#pragma unroll
        for (int i = 0; i < YS; i++) {
            const int k = tid + i * NT;
            if (k >= NF) break;
            for (int q = 0; q < 4; q++) sP[q * A2 + k] = yf[i][q];
        }
        __syncthreads();

        // 4. apply the fluxes (applyFluxes, example_Kelvin_Helmholtz/KH_DBA.cu:839-840)
        const double dtdx = c.dt * c.dx;
        double part[4] = {0.0, 0.0, 0.0, 0.0};
        double vfresh = 0.0;
        for (int k = tid; k < TB * TB; k += NT) {
            const int lx = k % TB, ly = k / TB;
            const int64_t gi = (x0 + lx) + (int64_t)(y0 + ly) * nx;
            const int fx = (lx + 1) + ly * FW, fxm = fx - 1;         // x-face arrays: (TB+1) x TB
            const int fy = lx + (ly + 1) * TB, fym = fy - TB;        // y-face arrays: TB x (TB+1)
            double un[4];
            for (int q = 0; q < 4; q++) {
                const double ch = dtdx * (-1.0 * sXF[q * NF + fx] + sXF[q * NF + fxm] - 1.0 * sP[q * A2 + fy] +
                                          sP[q * A2 + fym]);
                const double u0 = I.buf[pself][q][gi] - nb.shift[q][4];
                un[q] = u0 + ch;
                part[q] += ch;
                I.buf[pself ^ 1][q][gi] = un[q];
            }
            if (cfl) {
                double r, u, v, p;
                kh_prims<P2>(c, un[0], un[1], un[2], un[3], &r, &u, &v, &p);
                vfresh = fmax(vfresh, kh_speed(c, r, u, v, p));
            }
        }

        // 5. detector byte, per-tile speeds, conserved bookkeeping
        if (dba && tid == 0) I.nflag[0][t] = (uint8_t)any;
        if (cfl) {
            const double vs = block_max<NT>(vstale, sred);
            const double vf = block_max<NT>(vfresh, sred);
            if (tid == 0) {
                I.tspeed[t] = vf;
                I.tspeed[I.nt + t] = vs;
            }
        }
        for (int q = 0; q < ncons; q++) {
            const double tot = block_sum<NT>(part[q], sred);
            if (tid == 0) {
                I.tpart[(int64_t)q * I.nt + t] = tot;
                I.s_tile[((int64_t)q * 2 + (pself ^ 1)) * I.nt + t] = I.scal[8 + q];
            }
        }
        __syncthreads();
    }
}

// Per-tile wave speeds of the current state (all tiles), so the first step of
// a CFL run has its dt; the finalize then keeps them current.
template <int TB>
__global__ void kh_prime_kernel(const Inst* __restrict__ insts, int slot) {
    CInst& I = ((CInst*)insts)[slot];
    KhC c;
    c.g = I.gas_gamma;
    c.dx = I.dx;
    c.dy = I.dy;
    c.vol = I.vol;
    c.dt = 0.0;
    __shared__ double sred[5];
    for (int t = blockIdx.x; t < I.nt; t += gridDim.x) {
        const int x0 = (t % I.ntx) * TB, y0 = (t / I.ntx) * TB;
        const int pp = bit_of(I.par[0], t);
        double vm = 0.0;
        for (int k = threadIdx.x; k < TB * TB; k += blockDim.x) {
            const int64_t gi = (x0 + k % TB) + (int64_t)(y0 + k / TB) * I.nx;
            double sh[4];
            for (int q = 0; q < 4; q++)
                sh[q] = q < I.ncons ? I.scal[8 + q] - I.s_tile[((int64_t)q * 2 + pp) * I.nt + t] : 0.0;
            double r, u, v, p;
            kh_prims(c, I.buf[pp][0][gi] - sh[0], I.buf[pp][1][gi] - sh[1], I.buf[pp][2][gi] - sh[2],
                     I.buf[pp][3][gi] - sh[3], &r, &u, &v, &p);
            vm = fmax(vm, kh_speed(c, r, u, v, p));
        }
        vm = block_max<256>(vm, sred);
        if (threadIdx.x == 0) {
            I.tspeed[t] = vm;
            I.tspeed[I.nt + t] = vm;
        }
        __syncthreads();
    }
}

__global__ void kh_dt_kernel(const Inst* __restrict__ insts, int slot) {
    CInst& I = ((CInst*)insts)[slot];
    __shared__ double sred[5];
    double vm = 0.0;
    for (int t = threadIdx.x; t < I.nt; t += blockDim.x) vm = fmax(vm, I.tspeed[t]);
    vm = block_max<256>(vm, sred);
    if (threadIdx.x == 0 && I.dt_mode == 1) I.scal[1] = I.cfl * (I.dx / vm);
}

}  // namespace dba
