// k_ref3d.h - fused DBA steps of the reference's own 3D models on 4x4x4
// activation blocks (4x4x1 when nz == 1), one wave64 per block, periodic.
//
//  * Kobayashi-94 dendrite: Evolve_p(_wakeup) + Evolve_T(_wakeup) + the detector
//    of Wakeup_Next (example_Dendritic_growth/kernel.cu:181-293, 322-574):
//    split p / T sets; regular T = T + dt lapT + p_new - p, DBA T = (T + dt lapT)
//    + (p_new - p) only where p-awake; Pi = 3.14159 (kernel.cu:7).
//  * Ternary Flory-Huggins Model B with reaction: Calc_mu(_wakeup) + Update(_wakeup)
//    + copy-back + detector |lap z| > eps (example_Reactive_ModelB/kernel.cu:290-763),
//    conserved change capture (example_Reactive_ModelB/Conserved/kernel.cu:630-702).
//    The chemical potentials of sleeping neighbours are the ones they stored
//    when they last ran (a block stores its surface cells).
#pragma once
#include "common.h"

namespace dba {

struct NbTable3 {
    int tile[27];
    int par[2][27];
    int act[2][27];
    double shift[kMaxFields][27];
};

template <int TB, int TZ>
__device__ __forceinline__ void nb_fill3(NbTable3& nb, CInst& I, int x0, int y0, int z0, bool regular,
                                         int nsets, int ncons) {
    const int d = threadIdx.x;
    if (d >= 27) return;
    const int dx = d % 3 - 1, dy = (d / 3) % 3 - 1, dz = d / 9 - 1;
    const int rx = dx < 0 ? x0 - 1 : (dx > 0 ? x0 + TB : x0);
    const int ry = dy < 0 ? y0 - 1 : (dy > 0 ? y0 + TB : y0);
    const int rz = dz < 0 ? z0 - 1 : (dz > 0 ? z0 + TZ : z0);
    const int mx = map_index(rx, I.nx, 0), my = map_index(ry, I.ny, 0), mz = map_index(rz, I.nz, 0);
    const int tt = mx / TB + (my / TB) * I.ntx + (mz / TZ) * I.ntx * I.nty;
    nb.tile[d] = tt;
    for (int s = 0; s < nsets; s++) {
        nb.par[s][d] = bit_of(I.par[s], tt);
        nb.act[s][d] = regular ? 1 : bit_of(I.a_cur[s], tt);
    }
    for (int f = 0; f < kMaxFields; f++)
        nb.shift[f][d] = f < ncons ? (I.scal[8 + f] - I.s_tile[((int64_t)f * 2 + nb.par[0][d]) * I.nt + tt]) : 0.0;
}

// 3D wrap of a small periodic offset
__device__ __forceinline__ int wrapi(int i, int n) {
    while (i < 0) i += n;
    while (i >= n) i -= n;
    return i;
}

// 27-point isotropic Laplacian from an LDS box of extents WX x WY x WZ, in the
// summation order of example_Reactive_ModelB/kernel.cu:333-344.
template <int WX, int WY>
__device__ __forceinline__ double lds_lap27(const double* s, int x, int y, int z, double inv_dx2) {
#define S3(a, b, c) s[(x + (a)) + (y + (b)) * WX + (z + (c)) * WX * WY]
    return inv_dx2 *
           (-64.0 / 15.0 * S3(0, 0, 0) +
            7.0 / 15.0 * (S3(1, 0, 0) + S3(-1, 0, 0) + S3(0, 1, 0) + S3(0, -1, 0) + S3(0, 0, 1) + S3(0, 0, -1)) +
            0.1 * (S3(1, 1, 0) + S3(1, -1, 0) + S3(-1, 1, 0) + S3(-1, -1, 0) + S3(1, 0, 1) + S3(0, 1, 1) +
                   S3(0, -1, 1) + S3(-1, 0, 1) + S3(1, 0, -1) + S3(0, 1, -1) + S3(0, -1, -1) + S3(-1, 0, -1)) +
            1.0 / 30.0 * (S3(1, 1, 1) + S3(-1, 1, 1) + S3(-1, -1, 1) + S3(1, -1, 1) + S3(1, 1, -1) +
                          S3(-1, 1, -1) + S3(-1, -1, -1) + S3(1, -1, -1)));
#undef S3
}

template <int TB, int TZ, int H>
__device__ __forceinline__ int nb_dir3(int lx, int ly, int lz) {
    const int a = lx < H ? 0 : (lx < TB + H ? 1 : 2);
    const int b = ly < H ? 0 : (ly < TB + H ? 1 : 2);
    const int c = lz < H ? 0 : (lz < TZ + H ? 1 : 2);
    return a + 3 * b + 9 * c;
}

// ------------------------------------------------------------ Model B (ternary)
template <int TB, int TZ>
__global__ __launch_bounds__(64) void modelb_step_kernel(const WorkList wl) {
    constexpr int NT = 64;
    constexpr int W2X = TB + 4, W2Z = TZ + 4;  // x, y, z with a 2-cell halo
    constexpr int W1X = TB + 2, W1Z = TZ + 2;  // mu with a 1-cell halo
    constexpr int A2 = W2X * W2X * W2Z;
    constexpr int A1 = W1X * W1X * W1Z;
    __shared__ double sc[3 * A2];
    __shared__ double smu[3 * A1];
    __shared__ double sred[2];
    __shared__ int sflag[2];
    __shared__ NbTable3 nb;

    WorkCursor cur;
    work_begin(wl, cur, false);
    const int total = wl.offsets[wl.nseg];
    const int tid = threadIdx.x;
    for (;;) {
        CInst* Ip;
        int t;
        uint64_t rec;
        if (!work_next(wl, cur, total, &Ip, &t, &rec, false)) break;
        CInst& I = *Ip;
        const int nx = I.nx, ny = I.ny, nz = I.nz;
        const int x0 = (t % I.ntx) * TB, y0 = ((t / I.ntx) % I.nty) * TB, z0 = (t / (I.ntx * I.nty)) * TZ;
        const bool regular = I.mode == 0;
        const int ncons = I.ncons;
        const int64_t nxy = (int64_t)nx * ny;
        nb_fill3<TB, TZ>(nb, I, x0, y0, z0, regular, 1, ncons);
        __syncthreads();
        const int pself = nb.par[0][13];

        // x, y, z on the block + 2-cell halo
        for (int k = tid; k < A2; k += NT) {
            const int lx = k % W2X, ly = (k / W2X) % W2X, lz = k / (W2X * W2X);
            const int d = nb_dir3<TB, TZ, 2>(lx, ly, lz);
            const int64_t gi = wrapi(x0 + lx - 2, nx) + wrapi(y0 + ly - 2, ny) * (int64_t)nx +
                               wrapi(z0 + lz - 2, nz) * nxy;
            const int pp = nb.par[0][d];
            for (int f = 0; f < 3; f++) sc[f * A2 + k] = I.buf[pp][f][gi] - nb.shift[f][d];
        }
        __syncthreads();

        // chemical potentials on the block + 1-cell ring (Calc_mu_wakeup)
        int flag = 0;
        const double inv_dx2 = 1.0 / (I.dx * I.dx);
        for (int k = tid; k < A1; k += NT) {
            const int lx = k % W1X, ly = (k / W1X) % W1X, lz = k / (W1X * W1X);
            const int d = nb_dir3<TB, TZ, 1>(lx, ly, lz);
            const int cx = lx + 1, cy = ly + 1, cz = lz + 1;
            const int64_t gi = wrapi(x0 + lx - 1, nx) + wrapi(y0 + ly - 1, ny) * (int64_t)nx +
                               wrapi(z0 + lz - 1, nz) * nxy;
            if (!nb.act[0][d]) {
                for (int f = 0; f < 3; f++) smu[f * A1 + k] = I.aux[f][gi];
                continue;
            }
            const int q = cx + cy * W2X + cz * W2X * W2X;
            const double X = sc[q], Y = sc[A2 + q], Z = sc[2 * A2 + q];
            const double phi = I.buf[pself][3][gi];
            const double B = 1.0 - X - Y - Z;
            const double lx_ = lds_lap27<W2X, W2X>(sc, cx, cy, cz, inv_dx2);
            const double ly_ = lds_lap27<W2X, W2X>(sc + A2, cx, cy, cz, inv_dx2);
            const double lz_ = lds_lap27<W2X, W2X>(sc + 2 * A2, cx, cy, cz, inv_dx2);
            const double dfx = -1.0 + 1.0 / I.r_x + I.chi_xy * Y + I.chi_xz * Z + I.chi_xb * B - I.chi_xb * X -
                               I.chi_yb * Y - I.chi_zb * Z + log(X) / I.r_x - log(B) - I.epsx_sq * lx_;
            const double dfy = -1.0 + 1.0 / I.r_y + I.chi_xy * X + I.chi_yz * Z + I.chi_yb * B - I.chi_xb * X -
                               I.chi_yb * Y - I.chi_zb * Z + log(Y) / I.r_y - log(B) - I.epsy_sq * ly_;
            const double dfz = -1.0 + 1.0 / I.r_z + I.chi_xz * X + I.chi_yz * Y + I.chi_zb * B - I.chi_xb * X -
                               I.chi_yb * Y - I.chi_zb * Z + log(Z) / I.r_z - log(B) -
                               I.p_gel * phi * phi / 2.0 / (1.0 - I.z_crit) - I.epsz_sq * lz_ -
                               log(I.k_eq) / I.r_z;
            smu[k] = dfx;
            smu[A1 + k] = dfy;
            smu[2 * A1 + k] = dfz;
            if (d == 13) {
                flag |= fabs(lz_) > I.thresh_a;
                if (!regular) {
                    I.aux[0][gi] = dfx;
                    I.aux[1][gi] = dfy;
                    I.aux[2][gi] = dfz;
                }
            }
        }
        __syncthreads();

        // Update_wakeup on the block
        double part[3] = {0.0, 0.0, 0.0};
        for (int k = tid; k < TB * TB * TZ; k += NT) {
            const int lx = k % TB, ly = (k / TB) % TB, lz = k / (TB * TB);
            const int64_t gi = (x0 + lx) + (int64_t)(y0 + ly) * nx + (int64_t)(z0 + lz) * nxy;
            const int mq = (lx + 1) + (ly + 1) * W1X + (lz + 1) * W1X * W1X;
            const int cq = (lx + 2) + (ly + 2) * W2X + (lz + 2) * W2X * W2X;
            const double dfx = smu[mq], dfy = smu[A1 + mq], dfz = smu[2 * A1 + mq];
            const double R = I.k_0 * (exp(I.n_exp * I.v_x * dfx + I.m_exp * I.v_y * dfy) - exp(I.v_z * dfz));
            const double chx = I.dt * I.v_x *
                               (I.mob_x * lds_lap27<W1X, W1X>(smu, lx + 1, ly + 1, lz + 1, inv_dx2) - I.n_exp * R);
            const double chy = I.dt * I.v_y *
                               (I.mob_y * lds_lap27<W1X, W1X>(smu + A1, lx + 1, ly + 1, lz + 1, inv_dx2) -
                                I.m_exp * R);
            const double chz = I.dt * I.v_z *
                               (I.mob_z * lds_lap27<W1X, W1X>(smu + 2 * A1, lx + 1, ly + 1, lz + 1, inv_dx2) + R);
            I.buf[pself ^ 1][0][gi] = sc[cq] + chx;
            I.buf[pself ^ 1][1][gi] = sc[A2 + cq] + chy;
            I.buf[pself ^ 1][2][gi] = sc[2 * A2 + cq] + chz;
            part[0] += chx;
            part[1] += chy;
            part[2] += chz;
        }
        if (!regular) {
            const int any = block_any<NT>(flag, sflag);
            if (tid == 0) I.nflag[0][t] = (uint8_t)any;
        }
        for (int f = 0; f < ncons; f++) {
            const double tot = block_sum<NT>(part[f], sred);
            if (tid == 0) {
                I.tpart[(int64_t)f * I.nt + t] = tot;
                I.s_tile[((int64_t)f * 2 + (pself ^ 1)) * I.nt + t] = I.scal[8 + f];
            }
        }
        __syncthreads();
    }
}

// ------------------------------------------------------------ Kobayashi-94 dendrite
template <int TB, int TZ>
__global__ __launch_bounds__(64) void k94_step_kernel(const WorkList wl) {
    constexpr int NT = 64;
    constexpr int WX = TB + 2, WZ = TZ + 2;
    constexpr int A = WX * WX * WZ;
    __shared__ double sp[A];
    __shared__ double sT[A];
    __shared__ int sflag[2];
    __shared__ NbTable3 nb;

    WorkCursor cur;
    work_begin(wl, cur, false);
    const int total = wl.offsets[wl.nseg];
    const int tid = threadIdx.x;
    for (;;) {
        CInst* Ip;
        int t;
        uint64_t rec;
        if (!work_next(wl, cur, total, &Ip, &t, &rec, false)) break;
        CInst& I = *Ip;
        const int nx = I.nx, ny = I.ny, nz = I.nz;
        const int x0 = (t % I.ntx) * TB, y0 = ((t / I.ntx) % I.nty) * TB, z0 = (t / (I.ntx * I.nty)) * TZ;
        const bool regular = I.mode == 0;
        const int64_t nxy = (int64_t)nx * ny;
        nb_fill3<TB, TZ>(nb, I, x0, y0, z0, regular, 2, 0);
        __syncthreads();
        const int ap = nb.act[0][13], at = nb.act[1][13];
        const int pp0 = nb.par[0][13], pp1 = nb.par[1][13];
        for (int k = tid; k < A; k += NT) {
            const int lx = k % WX, ly = (k / WX) % WX, lz = k / (WX * WX);
            const int d = nb_dir3<TB, TZ, 1>(lx, ly, lz);
            const int64_t gi = wrapi(x0 + lx - 1, nx) + wrapi(y0 + ly - 1, ny) * (int64_t)nx +
                               wrapi(z0 + lz - 1, nz) * nxy;
            if (d == 13 || ap) sp[k] = I.buf[nb.par[0][d]][0][gi];
            if (d == 13 || at || regular) sT[k] = I.buf[nb.par[1][d]][1][gi];
        }
        __syncthreads();
        const double f0 = 1. / (2 * I.dx), f1 = 1. / I.dx / I.dx, f2 = -I.alpha / 3.14159, f3 = 1. / I.tau;
        const double dt = I.dt, eps = I.eps_bar, delta = I.delta, gam = I.gamma;
        int fp = 0, fT = 0;
        for (int k = tid; k < TB * TB * TZ; k += NT) {
            const int lx = k % TB, ly = (k / TB) % TB, lz = k / (TB * TB);
            const int64_t gi = (x0 + lx) + (int64_t)(y0 + ly) * nx + (int64_t)(z0 + lz) * nxy;
            const int q = (lx + 1) + (ly + 1) * WX + (lz + 1) * WX * WX;
            const double pc = sp[q], Tc = sT[q];
            double pn = pc;
            if (ap) {
                const double pxp = sp[q + 1], pxm = sp[q - 1], pyp = sp[q + WX], pym = sp[q - WX];
                const double pzp = sp[q + WX * WX], pzm = sp[q - WX * WX];
                const double dpx = (pxp - pxm) * f0, dpy = (pyp - pym) * f0, dpz = (pzp - pzm) * f0;
                const double lapP = (pxp + pxm + pyp + pym + pzp + pzm - 6 * pc) * f1;
                const double v2 = dpx * dpx + dpy * dpy + dpz * dpz;
                double sigma;
                if (v2 <= 1E-6)
                    sigma = 1;
                else
                    sigma = 1 - delta * (1 - ((dpx * dpx) * (dpx * dpx) + (dpy * dpy) * (dpy * dpy) +
                                              (dpz * dpz) * (dpz * dpz)) / (v2 * v2));
                const double m = f2 * atan(gam * sigma * Tc);
                const double p_change = f3 * (eps * eps * lapP + pc * (1 - pc) * (pc - 0.5 + m));
                pn = pc + dt * p_change;
                I.buf[pp0 ^ 1][0][gi] = pn;
                fp |= (pn >= I.thresh_a && pn <= (1.0 - I.thresh_a));
            }
            if (at) {
                const double lapT = (sT[q + 1] + sT[q - 1] + sT[q + WX] + sT[q - WX] + sT[q + WX * WX] +
                                     sT[q - WX * WX] - 6 * Tc) * f1;
                double tn;
                if (regular) {
                    tn = Tc + dt * lapT + pn - pc;
                } else {
                    tn = Tc + dt * lapT;
                    if (ap) tn += pn - pc;
                }
                I.buf[pp1 ^ 1][1][gi] = tn;
                fT |= fabs(tn - Tc) >= I.thresh_b;
            }
        }
        if (!regular) {
            const int anyp = block_any<NT>(fp, sflag);
            const int anyT = block_any<NT>(fT, sflag);
            if (tid == 0) {
                if (ap) I.nflag[0][t] = (uint8_t)anyp;
                if (at) I.nflag[1][t] = (uint8_t)anyT;
            }
        }
        __syncthreads();
    }
}

}  // namespace dba
