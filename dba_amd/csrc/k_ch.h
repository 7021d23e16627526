// k_ch.h - fused DBA step for 2D Cahn-Hilliard with a linear reaction:
//     dc/dt = M lap(mu) - Gamma (c - cbar),   mu = c^3 - c - kappa lap c.
//
// Reference counterpart: the two passes Calc_mu_wakeup -> Update_wakeup, the
// copy-back Swap_wakeup and the detector of Wakeup_Next
// (example_Reactive_ModelB/kernel.cu:519-654, 469-517, 657-713), plus the
// conserved change capture (example_Reactive_ModelB/Conserved/kernel.cu:630-702).
//
// One workgroup per active tile per worklist item:
//   1. stage c on the tile plus a 2-cell halo in LDS (coalesced rows; halo cells
//      come from each owner tile's current buffer, minus its conserved shift);
//   2. mu on the tile plus a 1-cell ring: computed here for the tile and for ring
//      cells of awake neighbours (what their own pass 1 computes this step); ring
//      cells of sleeping neighbours use the mu they stored when they last ran,
//      which is what the reference's full mu array holds for them;
//   3. c' = c + dt (M lap mu - Gamma (c - cbar)) into the other buffer;
//   4. detector any(|lap c| > eps) -> atomicOr dilation into the next bitmap;
//      the tile's own edge-ring mu is stored for its neighbours.
// HBM traffic per cell: read c, write c' (16 B) + the mu ring (4(T-1)/T^2 of cells).
#pragma once
#include "common.h"

namespace dba {

template <int W>
__device__ __forceinline__ double lds_lap5(const double* s, int x, int y, double invdx2) {
    double s1 = (s[(x + 1) + y * W] + s[(x - 1) + y * W]) + (s[x + (y + 1) * W] + s[x + (y - 1) * W]);
    return (s1 - 4.0 * s[x + y * W]) * invdx2;
}

template <int W>
__device__ __forceinline__ double lds_lap9(const double* s, int x, int y, double inv6dx2) {
    double s1 = (s[(x + 1) + y * W] + s[(x - 1) + y * W]) + (s[x + (y + 1) * W] + s[x + (y - 1) * W]);
    double s2 = (s[(x + 1) + (y + 1) * W] + s[(x - 1) + (y + 1) * W]) +
                (s[(x + 1) + (y - 1) * W] + s[(x - 1) + (y - 1) * W]);
    return ((4.0 * s1 + s2) - 20.0 * s[x + y * W]) * inv6dx2;
}

template <int TB, int NT>
__global__ __launch_bounds__(NT) void ch_step_kernel(const Inst* __restrict__ insts,
                                                    const int32_t* __restrict__ offsets, int n_inst) {
    constexpr int W2 = TB + 4;
    constexpr int W1 = TB + 2;
    __shared__ double sc[W2 * W2];
    __shared__ double smu[W1 * W1];
    __shared__ double sred[NT / 64 + 1];
    __shared__ int sflag[NT / 64 + 1];

    const int total = offsets[n_inst];
    const int G = gridDim.x;
    const int tid = threadIdx.x;
    for (int item = xcd_remap(blockIdx.x, G); item < total; item += G) {
        const int slot = find_slot(offsets, n_inst, item);
        const Inst& I = insts[slot];
        const int t = I.seg[item - offsets[slot]];
        const int nx = I.nx, ny = I.ny;
        const int tix = t % I.ntx, tiy = t / I.ntx;
        const int x0 = tix * TB, y0 = tiy * TB;
        const bool regular = I.mode == 0;
        const bool cons = I.mode == 2;
        const int pself = bit_of(I.par[0], t);
        const double* cur = I.buf[pself][0];
        double* nxt = I.buf[pself ^ 1][0];
        const double S = cons ? I.scal[8] : 0.0;
        const double shift_self = cons ? (S - I.s_tile[pself * I.nt + t]) : 0.0;
        const bool nine = I.stencil != 5;

        // 1. c on the tile + 2-cell halo
        for (int k = tid; k < W2 * W2; k += NT) {
            const int lx = k % W2, ly = k / W2;
            const int gx = x0 + lx - 2, gy = y0 + ly - 2;
            double v;
            if (lx >= 2 && lx < TB + 2 && ly >= 2 && ly < TB + 2) {
                v = cur[gx + (int64_t)gy * nx] - shift_self;
            } else {
                const int mx = map_index(gx, nx, I.bc_x), my = map_index(gy, ny, I.bc_y);
                const int tt = mx / TB + (my / TB) * I.ntx;
                const int pp = bit_of(I.par[0], tt);
                v = I.buf[pp][0][mx + (int64_t)my * nx];
                if (cons) v -= (S - I.s_tile[pp * I.nt + tt]);
            }
            sc[k] = v;
        }
        __syncthreads();

        // 2. mu on the tile + 1-cell ring; detector on the tile
        int flag = 0;
        for (int k = tid; k < W1 * W1; k += NT) {
            const int lx = k % W1, ly = k / W1;
            const bool self = lx >= 1 && lx <= TB && ly >= 1 && ly <= TB;
            bool fresh = self || regular;
            int64_t gidx = 0;
            if (!self) {
                const int mx = map_index(x0 + lx - 1, nx, I.bc_x), my = map_index(y0 + ly - 1, ny, I.bc_y);
                const int tt = mx / TB + (my / TB) * I.ntx;
                gidx = mx + (int64_t)my * nx;
                if (!fresh) fresh = (tt == t) || bit_of(I.a_cur[0], tt);
            }
            double mu;
            if (fresh) {
                const double lc = nine ? lds_lap9<W2>(sc, lx + 1, ly + 1, I.inv6dx2)
                                       : lds_lap5<W2>(sc, lx + 1, ly + 1, I.invdx2);
                const double cv = sc[(lx + 1) + (ly + 1) * W2];
                mu = (cv * cv * cv - cv) - I.kappa * lc;
                if (self) {
                    flag |= fabs(lc) > I.thresh_a;
                    if (!regular && (lx == 1 || lx == TB || ly == 1 || ly == TB))
                        I.aux[0][(x0 + lx - 1) + (int64_t)(y0 + ly - 1) * nx] = mu;
                }
            } else {
                mu = I.aux[0][gidx];
            }
            smu[k] = mu;
        }
        __syncthreads();

        // 3. c' on the tile
        double part = 0.0;
        for (int k = tid; k < TB * TB; k += NT) {
            const int lx = k % TB, ly = k / TB;
            const double lm = nine ? lds_lap9<W1>(smu, lx + 1, ly + 1, I.inv6dx2)
                                   : lds_lap5<W1>(smu, lx + 1, ly + 1, I.invdx2);
            const double cv = sc[(lx + 2) + (ly + 2) * W2];
            const double change = I.dt * (I.mobility * lm - I.react_gamma * (cv - I.cbar));
            nxt[(x0 + lx) + (int64_t)(y0 + ly) * nx] = cv + change;
            part += change;
        }

        // 4. detector + dilation, conserved bookkeeping
        if (!regular) {
            const int any = block_any<NT>(flag, sflag);
            if (tid == 0 && any) dilate_into(I.a_next[0], I, t, 0);
        }
        if (cons) {
            const double tot = block_sum<NT>(part, sred);
            if (tid == 0) {
                I.tpart[t] = tot;
                I.s_tile[(pself ^ 1) * I.nt + t] = S;
            }
        }
        __syncthreads();
    }
}

}  // namespace dba
