// k_k93.h - fused DBA step for the 2D Kobayashi-93 anisotropic dendrite:
//   tau dphi/dt = -d/dx(eps eps' phi_y) + d/dy(eps eps' phi_x) + eps^2 lap phi
//                 + phi (1 - phi)(phi - 1/2 + m),   m = (alpha/pi) atan(gamma (T_eq - T)),
//   dT/dt = lap T + K dphi/dt,   eps = eps_bar (1 + delta cos j(theta - theta0)).
//
// Reference counterpart: Evolve_p_wakeup, Evolve_T_wakeup and the detector of
// Wakeup_Next (example_Dendritic_growth/kernel.cu:322-574), with the same split
// p / T activation sets: phi advances on p-awake tiles, T on T-awake tiles, the
// latent-heat term is added only where the tile is p-awake, and the detectors
// are Metric_p <= phi' <= 1 - Metric_p and |T' - T| >= Metric_T.
//
// cos/sin of j*theta come from the j-th power of (phi_x + i phi_y)/|grad phi|,
// so the angle needs no atan2/cos/sin (theta = 0 where grad phi = 0).
// One workgroup per tile: phi staged with a 2-cell halo (the cross terms need
// eps eps' grad phi on a 1-cell ring), T with a 1-cell halo.
#pragma once
#include "common.h"

namespace dba {

__device__ __forceinline__ void k93_aniso(const Inst& I, double gx, double gy, double* eps, double* a) {
    double g2 = gx * gx + gy * gy;
    double cj = 1.0, sj = 0.0;
    if (g2 != 0.0) {
        double r = sqrt(g2);
        double u = gx / r, v = gy / r;
        double cr = u, ci = v;
        for (int k = 1; k < I.aniso_j; k++) {
            double tt = cr * u - ci * v;
            ci = cr * v + ci * u;
            cr = tt;
        }
        cj = cr;
        sj = ci;
    }
    double ca = cj * I.cos_jt0 + sj * I.sin_jt0;
    double sa = sj * I.cos_jt0 - cj * I.sin_jt0;
    double e = I.eps_bar * (1.0 + I.delta * ca);
    double ep = I.k_epsp * sa;
    *eps = e;
    *a = e * ep;
}

template <int TB, int NT>
__global__ __launch_bounds__(NT) void k93_step_kernel(const Inst* __restrict__ insts,
                                                     const int32_t* __restrict__ offsets, int n_inst) {
    constexpr int W2 = TB + 4;
    constexpr int W1 = TB + 2;
    __shared__ double sp[W2 * W2];
    __shared__ double sT[W1 * W1];
    __shared__ double sFx[W1 * W1];
    __shared__ double sFy[W1 * W1];
    __shared__ double sE2[TB * TB];
    __shared__ int sflag[NT / 64 + 1];

    const int total = offsets[n_inst];
    const int G = gridDim.x;
    const int tid = threadIdx.x;
    for (int item = xcd_remap(blockIdx.x, G); item < total; item += G) {
        const int slot = find_slot(offsets, n_inst, item);
        const Inst& I = insts[slot];
        const int t = I.seg[item - offsets[slot]];
        const int nx = I.nx, ny = I.ny;
        const int x0 = (t % I.ntx) * TB, y0 = (t / I.ntx) * TB;
        const bool regular = I.mode == 0;
        const int ap = regular ? 1 : bit_of(I.a_cur[0], t);
        const int at = regular ? 1 : bit_of(I.a_cur[1], t);
        const int pp0 = bit_of(I.par[0], t);
        const int pp1 = bit_of(I.par[1], t);

        // stage phi (2-cell halo) when p-awake; T (1-cell halo) when T-awake,
        // and phi / T of the tile itself always (the T update needs phi - phi')
        for (int k = tid; k < W2 * W2; k += NT) {
            const int lx = k % W2, ly = k / W2;
            const bool self = lx >= 2 && lx < TB + 2 && ly >= 2 && ly < TB + 2;
            if (!self && !ap) continue;
            const int gx = x0 + lx - 2, gy = y0 + ly - 2;
            double v;
            if (self) {
                v = I.buf[pp0][0][gx + (int64_t)gy * nx];
            } else {
                const int mx = map_index(gx, nx, I.bc_x), my = map_index(gy, ny, I.bc_y);
                const int tt = mx / TB + (my / TB) * I.ntx;
                v = I.buf[bit_of(I.par[0], tt)][0][mx + (int64_t)my * nx];
            }
            sp[k] = v;
        }
        for (int k = tid; k < W1 * W1; k += NT) {
            const int lx = k % W1, ly = k / W1;
            const bool self = lx >= 1 && lx <= TB && ly >= 1 && ly <= TB;
            if (!self && !at) continue;
            const int gx = x0 + lx - 1, gy = y0 + ly - 1;
            double v;
            if (self) {
                v = I.buf[pp1][1][gx + (int64_t)gy * nx];
            } else {
                const int mx = map_index(gx, nx, I.bc_x), my = map_index(gy, ny, I.bc_y);
                const int tt = mx / TB + (my / TB) * I.ntx;
                v = I.buf[bit_of(I.par[1], tt)][1][mx + (int64_t)my * nx];
            }
            sT[k] = v;
        }
        __syncthreads();

        // eps eps' grad phi on the tile + ring, eps^2 on the tile
        if (ap) {
            for (int k = tid; k < W1 * W1; k += NT) {
                const int lx = k % W1, ly = k / W1;
                const int cx = lx + 1, cy = ly + 1;
                const double gx = (sp[(cx + 1) + cy * W2] - sp[(cx - 1) + cy * W2]) * I.inv2dx;
                const double gy = (sp[cx + (cy + 1) * W2] - sp[cx + (cy - 1) * W2]) * I.inv2dx;
                double e, a;
                k93_aniso(I, gx, gy, &e, &a);
                sFx[k] = a * gx;
                sFy[k] = a * gy;
                if (lx >= 1 && lx <= TB && ly >= 1 && ly <= TB) sE2[(lx - 1) + (ly - 1) * TB] = e * e;
            }
        }
        __syncthreads();

        int fp = 0, fT = 0;
        for (int k = tid; k < TB * TB; k += NT) {
            const int lx = k % TB, ly = k / TB;
            const int64_t gi = (x0 + lx) + (int64_t)(y0 + ly) * nx;
            const double ph = sp[(lx + 2) + (ly + 2) * W2];
            const double Tv = sT[(lx + 1) + (ly + 1) * W1];
            double pn = ph;
            if (ap) {
                const int cx = lx + 2, cy = ly + 2, fx = lx + 1, fy = ly + 1;
                const double cross = ((sFx[fx + (fy + 1) * W1] - sFx[fx + (fy - 1) * W1]) -
                                      (sFy[(fx + 1) + fy * W1] - sFy[(fx - 1) + fy * W1])) * I.inv2dx;
                const double lap = lds_lap5<W2>(sp, cx, cy, I.invdx2);
                const double m = I.alpha_pi * atan(I.gamma * (I.t_eq - Tv));
                const double rhs = (cross + sE2[k] * lap) + (ph * (1.0 - ph)) * ((ph - 0.5) + m);
                pn = ph + I.dt_tau * rhs;
                I.buf[pp0 ^ 1][0][gi] = pn;
                fp |= (pn >= I.thresh_a && pn <= (1.0 - I.thresh_a));
            }
            if (at) {
                const double lapT = lds_lap5<W1>(sT, lx + 1, ly + 1, I.invdx2);
                const double t1 = Tv + I.dt * lapT;
                const double Tn = ap ? t1 + I.latent_k * (pn - ph) : t1;
                I.buf[pp1 ^ 1][1][gi] = Tn;
                fT |= fabs(Tn - Tv) >= I.thresh_b;
            }
        }

        if (!regular) {
            const int anyp = block_any<NT>(fp, sflag);
            const int anyT = block_any<NT>(fT, sflag);
            if (tid == 0) {
                if (ap && anyp) dilate_into(I.a_next[0], I, t, 0);
                if (at && anyT) dilate_into(I.a_next[1], I, t, 0);
            }
        }
        __syncthreads();
    }
}

}  // namespace dba
