// k_ch.h - fused DBA step for 2D Cahn-Hilliard with a linear reaction:
//     dc/dt = M lap(mu) - Gamma (c - cbar),   mu = c^3 - c - kappa lap c.
//
// Reference counterpart: the two passes Calc_mu_wakeup -> Update_wakeup, the
// copy-back Swap_wakeup and the detector of Wakeup_Next
// (example_Reactive_ModelB/kernel.cu:519-654, 469-517, 657-713), plus the
// conserved change capture (example_Reactive_ModelB/Conserved/kernel.cu:630-702).
//
// One workgroup per awake tile (worklist item):
//   0. resolve the 3x3 neighbour tiles once (owner tile, buffer parity, awake,
//      conserved shift) into LDS;
//   1. stage c on the tile plus a 2-cell halo in LDS (coalesced rows; halo cells
//      come from each owner tile's current buffer, minus its conserved shift);
//   2. mu on the tile plus a 1-cell ring: computed here for the tile and for ring
//      cells of awake neighbours (what their own pass 1 computes this step); ring
//      cells of sleeping neighbours use the mu they stored when they last ran,
//      which is what the reference's full mu array holds for them;
//   3. c' = c + dt (M lap mu - Gamma (c - cbar)) into the other buffer;
//   4. detector any(|lap c| > eps) -> one byte per tile; the tile's own
//      edge-ring mu is stored for its neighbours.
// HBM traffic per cell: read c, write c' (16 B) + the mu ring (4(T-1)/T^2 of cells),
// kept as one contiguous record of 4 TB - 4 values per tile (ring order of
// TileRing) so that its column cells do not each cost a partial-line write.
#pragma once
#include "common.h"
#include "k_finalize.h"

namespace dba {

// Ring cells of a TB x TB tile, in record order: bottom row, top row, left
// column, right column (the columns without their corners).
template <int TB>
struct TileRing {
    static constexpr int N = 4 * TB - 4;
    __device__ __forceinline__ static int slot(int lx, int ly) {
        if (ly == 0) return lx;
        if (ly == TB - 1) return TB + lx;
        if (lx == 0) return 2 * TB + ly - 1;
        if (lx == TB - 1) return 3 * TB - 2 + ly - 1;
        return -1;
    }
};

// Column-strip register blocking: thread (tx, ty) owns rows ty*R .. ty*R+R-1
// of column tx, R = TB*TB/NT. A 9-point stencil over its R cells reads
// 3 (R + 2) LDS values instead of 9 R.
template <int W, int R, bool NINE>
__device__ __forceinline__ void strip_lap(const double* s, int cx, int cy0, double inv, double* out) {
    double L[R + 2], C[R + 2], Rt[R + 2];
#pragma unroll
    for (int r = 0; r < R + 2; r++) {
        const int row = (cy0 - 1 + r) * W;
        L[r] = s[row + cx - 1];
        C[r] = s[row + cx];
        Rt[r] = s[row + cx + 1];
    }
#pragma unroll
    for (int r = 0; r < R; r++) {
        const double s1 = (Rt[r + 1] + L[r + 1]) + (C[r + 2] + C[r]);
        if (NINE) {
            const double s2 = (Rt[r + 2] + L[r + 2]) + (Rt[r] + L[r]);
            out[r] = ((4.0 * s1 + s2) - 20.0 * C[r + 1]) * inv;
        } else {
            out[r] = (s1 - 4.0 * C[r + 1]) * inv;
        }
    }
}

// One tile's step. WORKLIST: fill the neighbour table from the record and store
// the detector byte here; otherwise the caller has filled the table (then a
// barrier) and stores the returned detector bit any(|lap c| > eps) (0 in regular
// mode, uniform over the workgroup) itself.
template <int TB, int NT, bool NINE, bool WORKLIST>
__device__ __forceinline__ int ch_tile(CInst& I, int t, uint64_t rec, NbTable& nb, double* sc, double* smu,
                                       double* sred, int* sflag) {
    constexpr int W2 = TB + 4;
    constexpr int W1 = TB + 2;
    constexpr int R = TB * TB / NT;
    constexpr int NH2 = W2 * W2 - TB * TB;  // halo cells of the c region
    constexpr int NH1 = W1 * W1 - TB * TB;  // ring cells of the mu region
    const int tid = threadIdx.x;
    const int tx = tid % TB, ty = tid / TB;
    const int nx = I.nx, ny = I.ny, ntx = I.ntx, bcx = I.bc_x, bcy = I.bc_y;
    const int x0 = (t % ntx) * TB, y0 = (t / ntx) * TB;
    const bool regular = I.mode == 0;
    const bool cons = I.mode == 2;
    const gdouble* __restrict__ b0 = I.buf[0][0];
    const gdouble* __restrict__ b1 = I.buf[1][0];
    gdouble* __restrict__ mu_store = I.aux[0];
    const double kappa = I.kappa, eps = I.thresh_a;
    const double inv = NINE ? I.inv6dx2 : I.invdx2;
    if (WORKLIST) {
        nb_fill<TB>(nb, I, t, x0, y0, regular, 1, cons ? 1 : 0, rec);
        __syncthreads();
    }
    const int pself = nb.par[0][4];
    const double sh_self = nb.shift[0][4];
    const gdouble* __restrict__ cur = pself ? b1 : b0;
    gdouble* __restrict__ nxt = pself ? I.buf[0][0] : I.buf[1][0];

    // 1. c: the thread's strip from HBM into registers and LDS, then the halo
    double c[R];
    const int gx = x0 + tx;
#pragma unroll
    for (int r = 0; r < R; r++) {
        c[r] = cur[gx + (int64_t)(y0 + ty * R + r) * nx] - sh_self;
        sc[(tx + 2) + (ty * R + r + 2) * W2] = c[r];
    }
    for (int k = tid; k < NH2; k += NT) {
        int lx, ly;
        if (k < 2 * W2) { ly = k / W2; lx = k % W2; }                      // two rows below
        else if (k < 4 * W2) { ly = TB + 2 + (k - 2 * W2) / W2; lx = (k - 2 * W2) % W2; }  // two rows above
        else { const int q = k - 4 * W2; ly = 2 + q / 4; const int c4 = q % 4; lx = c4 < 2 ? c4 : TB + c4; }
        const int d = nb_dir1<TB, 2>(lx) + 3 * nb_dir1<TB, 2>(ly);
        const int mx = map_index(x0 + lx - 2, nx, bcx), my = map_index(y0 + ly - 2, ny, bcy);
        sc[lx + ly * W2] = (nb.par[0][d] ? b1 : b0)[mx + (int64_t)my * nx] - nb.shift[0][d];
    }
    __syncthreads();

    // 2. mu on the strip (+ detector), then on the 1-cell ring
    int flag = 0;
    {
        double lc[R];
        strip_lap<W2, R, NINE>(sc, tx + 2, ty * R + 2, inv, lc);
#pragma unroll
        for (int r = 0; r < R; r++) {
            const double cv = c[r];
            const double mu = (cv * cv * cv - cv) - kappa * lc[r];
            flag |= fabs(lc[r]) > eps;
            const int ly = ty * R + r;
            smu[(tx + 1) + (ly + 1) * W1] = mu;
            const int slot = TileRing<TB>::slot(tx, ly);
            if (!regular && slot >= 0) mu_store[(int64_t)t * TileRing<TB>::N + slot] = mu;
        }
    }
    for (int k = tid; k < NH1; k += NT) {
        int lx, ly;
        if (k < W1) { ly = 0; lx = k; }
        else if (k < 2 * W1) { ly = TB + 1; lx = k - W1; }
        else { const int q = k - 2 * W1; ly = 1 + q / 2; lx = (q % 2) ? TB + 1 : 0; }
        const int d = nb_dir1<TB, 1>(lx) + 3 * nb_dir1<TB, 1>(ly);
        double mu;
        if (nb.act[0][d]) {
            const double lcv = NINE ? lds_lap9<W2>(sc, lx + 1, ly + 1, inv) : lds_lap5<W2>(sc, lx + 1, ly + 1, inv);
            const double cv = sc[(lx + 1) + (ly + 1) * W2];
            mu = (cv * cv * cv - cv) - kappa * lcv;
        } else {
            // the ring value the sleeping owner stored (a ring cell of it: the map
            // only lands inside the running tile itself)
            const int mx = map_index(x0 + lx - 1, nx, bcx), my = map_index(y0 + ly - 1, ny, bcy);
            const int sl = TileRing<TB>::slot(mx & (TB - 1), my & (TB - 1));
            mu = sl >= 0 ? mu_store[(int64_t)nb.tile[d] * TileRing<TB>::N + sl] : 0.0;
        }
        smu[lx + ly * W1] = mu;
    }
    __syncthreads();

    // 3. c' on the strip
    double part = 0.0;
    {
        double lm[R];
        strip_lap<W1, R, NINE>(smu, tx + 1, ty * R + 1, inv, lm);
        const double dt = I.dt, mob = I.mobility, rg = I.react_gamma, cbar = I.cbar;
#pragma unroll
        for (int r = 0; r < R; r++) {
            const double change = dt * (mob * lm[r] - rg * (c[r] - cbar));
            nxt[gx + (int64_t)(y0 + ty * R + r) * nx] = c[r] + change;
            part += change;
        }
    }

    // 4. detector bit, conserved bookkeeping
    const int any = regular ? 0 : block_any<NT>(flag, sflag);
    if (WORKLIST && !regular && tid == 0) I.nflag[0][t] = (uint8_t)any;
    if (cons) {
        const double tot = block_sum<NT>(part, sred);
        if (tid == 0) {
            if (WORKLIST) {
                I.tpart[t] = tot;
                I.s_tile[(pself ^ 1) * I.nt + t] = I.scal[8];
            } else {
                nb.tot = tot;  // the single-launch caller does the bookkeeping
            }
        }
    }
    __syncthreads();
    return any;
}

template <int TB, int NT>
__global__ __launch_bounds__(NT) void ch_step_kernel(const WorkList wl) {
    static_assert(NT % TB == 0 && TB * TB % NT == 0, "thread layout needs TB | NT | TB*TB");
    constexpr int W2 = TB + 4;
    constexpr int W1 = TB + 2;
    __shared__ double sc[W2 * W2];
    __shared__ double smu[W1 * W1];
    __shared__ double sred[NT / 64 + 1];
    __shared__ int sflag[NT / 64 + 1];
    __shared__ NbTable nb;

    WorkCursor cur;
    work_begin(wl, cur, true);
    const int total = wl.offsets[wl.nseg];
    for (;;) {
        CInst* Ip;
        int t;
        uint64_t rec;
        if (!work_next(wl, cur, total, &Ip, &t, &rec)) break;
        CInst& I = *Ip;
        if (I.stencil != 5) ch_tile<TB, NT, true, true>(I, t, rec, nb, sc, smu, sred, sflag);
        else ch_tile<TB, NT, false, true>(I, t, rec, nb, sc, smu, sred, sflag);
    }
}

// Single-launch DBA step of a CH group: one workgroup per tile of every
// instance and no worklist or finalize launch. Workgroup t derives the awake bit and
// buffer parity of its 3x3 owner tiles itself: at the first step of a run from
// the bitmaps, afterwards from the per-tile bytes the previous step wrote
// (awake = the dilation of the detector bits, exactly the finalize's rule). A
// sleeping tile writes its bytes and exits; an awake one runs ch_tile as in the
// worklist kernel. The canonicalising kernels (dense_canon_*) restore the
// bitmaps and counters at the end of the run. Results are bit-identical to the
// worklist path.
template <int TB, int NT, bool CONS>
__global__ __launch_bounds__(NT) void ch_dense_kernel(const Inst* __restrict__ insts,
                                                      const int32_t* __restrict__ dmap, int first, int s) {
    constexpr int W2 = TB + 4;
    constexpr int W1 = TB + 2;
    __shared__ double sc[W2 * W2];
    __shared__ double smu[W1 * W1];
    __shared__ double sred[NT / 64 + 1];
    __shared__ int sflag[NT / 64 + 1];
    __shared__ NbTable nb;
    __shared__ double sW[kFinNT / 64];
    int t;
    CInst& I = dense_item(insts, dmap, &t);
    const int prev = s & 1, next = prev ^ 1;
    const bool regular = I.mode == 0;
    constexpr bool cons = CONS;  // conserved DBA: one instance in one segment (the host checks)
    dense_nb_fill(nb, I, t, first, prev, regular, 1, cons);
    if (cons && !first) {
        // The previous step's change sums (tpart + (1 + parity) nt, 0 for tiles
        // that did not run) in the finalize's order: 16 words of 64 tiles, each
        // by the shuffle tree of block_sum, then the words in order.
        constexpr int NWV = NT / 64, PER = (kFinNT / 64 + NWV - 1) / NWV;
        const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
        const gdouble* tp = I.tpart + (int64_t)(1 + prev) * I.nt;
        double v[PER];
#pragma unroll
        for (int k = 0; k < PER; k++) {  // all of this wave's words in flight at once
            const int w = wid + k * NWV, tt = w * 64 + lane;
            v[k] = (w < kFinNT / 64 && tt < I.nt) ? tp[tt] : 0.0;
        }
#pragma unroll
        for (int k = 0; k < PER; k++) {
            v[k] = wave_sum_down(v[k]);  // block_sum's tree
            const int w = wid + k * NWV;
            if (lane == 0 && w < kFinNT / 64) sW[w] = v[k];
        }
    }
    __syncthreads();
    if (cons) {
        if (threadIdx.x == 0) {
            // the cumulative shift of this step, as the finalize advances it
            double S = I.scal[8];
            if (!first) {
                double r = 0.0;
                for (int w = 0; w < kFinNT / 64; w++) r += sW[w];
                S = I.scal[12 + prev] + r / I.ntotal;
            }
            nb.S = S;
            for (int d = 0; d < 9; d++) nb.shift[0][d] = S - nb.shift[0][d];  // epoch -> pending shift
            if (t == 0) I.scal[12 + next] = S;
        }
        __syncthreads();
    }
    const int pself = nb.par[0][4], ran = nb.act[0][4];
    int any = 0;
    if (ran)
        any = I.stencil != 5 ? ch_tile<TB, NT, true, false>(I, t, 0, nb, sc, smu, sred, sflag)
                             : ch_tile<TB, NT, false, false>(I, t, 0, nb, sc, smu, sred, sflag);
    if (threadIdx.x == 0) {
        I.dF[next][t] = (uint8_t)any;
        I.dP[next][t] = (uint8_t)(pself ^ ran);
        if (ran) {
            I.druns[t] += 1;
            const int64_t step = I.ctr[0] + s;
            if (step < I.hist_cap) atomicAdd((int*)&I.hist[2 * step], 1);
        }
        if (cons) {
            I.tpart[(int64_t)(1 + next) * I.nt + t] = ran ? nb.tot : 0.0;
            if (ran) I.s_tile[(pself ^ 1) * I.nt + t] = nb.S;
        }
    }
}

}  // namespace dba
