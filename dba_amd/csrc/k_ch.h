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
//   4. detector any(|lap c| > eps) -> one byte per tile; a tile that did not
//      fire stores its own edge-ring mu for its neighbours (one that fired is
//      awake next step, so nobody reads that step's ring).
// HBM traffic per cell: read c, write c' (16 B) + the mu ring (4(T-1)/T^2 of cells)
// of tiles that did not fire, kept as one contiguous record of 4 TB - 4 values
// per tile (ring order of TileRing) so that its column cells do not each cost a
// partial-line write.
#pragma once
#include "common.h"
#include "k_finalize.h"

namespace dba {

// Ring cells of a w x h tile (TB x TB, or smaller for a partial last tile), in
// record order: bottom row, top row, left column, right column (the columns
// without their corners). A tile's record holds N = 4 TB - 4 slots; a partial
// tile uses the first 2 w + 2 h - 4 of them.
template <int TB>
struct TileRing {
    static constexpr int N = 4 * TB - 4;
    __device__ __forceinline__ static int slot(int lx, int ly, int w = TB, int h = TB) {
        if (ly == 0) return lx;
        if (ly == h - 1) return w + lx;
        if (lx == 0) return 2 * w + ly - 1;
        if (lx == w - 1) return 2 * w + h - 2 + ly - 1;
        return -1;
    }
    // the cell of slot k (the inverse of slot)
    __device__ __forceinline__ static void cell(int k, int* lx, int* ly, int w = TB, int h = TB) {
        if (k < w) { *lx = k; *ly = 0; }
        else if (k < 2 * w) { *lx = k - w; *ly = h - 1; }
        else if (k < 2 * w + h - 2) { *lx = 0; *ly = k - 2 * w + 1; }
        else { *lx = w - 1; *ly = k - (2 * w + h - 2) + 1; }
    }
    __device__ __forceinline__ static int size(int w = TB, int h = TB) { return 2 * w + 2 * h - 4; }
    // the stored ring value of cell (mx, my) of its sleeping owner tile
    __device__ __forceinline__ static double stale(const gdouble* store, int owner, int mx, int my, int nx, int ny) {
        const int ox = mx & ~(TB - 1), oy = my & ~(TB - 1);
        const int sl = slot(mx - ox, my - oy, min(TB, nx - ox), min(TB, ny - oy));
        return sl >= 0 ? store[(int64_t)owner * N + sl] : 0.0;
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

// One partial tile's step (the last tile column / row when the grid is not a
// multiple of the tile): w x h cells, a generic cell loop instead of the
// register strips, and the owner of every halo cell found from its mapped
// position (past a partial tile, the offset in the tile's region no longer
// tells which neighbour holds a cell). The same arithmetic as ch_tile.
template <int TB, int NT, bool NINE, bool WORKLIST>
__device__ __noinline__ int ch_tile_partial(CInst& I, int t, uint64_t rec, NbTable& nb, double* sc, double* smu,
                                            double* sred, int* sflag) {
    constexpr int W2 = TB + 4;
    constexpr int W1 = TB + 2;
    const int tid = threadIdx.x;
    const int nx = I.nx, ny = I.ny, ntx = I.ntx, nty = I.nty, bcx = I.bc_x, bcy = I.bc_y;
    const int bcyo = bcy == 0 ? 0 : 1;
    const int ti = t % ntx, tj = t / ntx;
    const int x0 = ti * TB, y0 = tj * TB;
    const int w = tile_extent(ti, TB, nx), h = tile_extent(tj, TB, ny);
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
    gdouble* __restrict__ nxt = pself ? I.buf[0][0] : I.buf[1][0];
    // 1. c on the tile and a 2-cell halo
    for (int k = tid; k < (w + 4) * (h + 4); k += NT) {
        const int lx = k % (w + 4), ly = k / (w + 4);
        const int mx = map_index(x0 + lx - 2, nx, bcx), my = map_index(y0 + ly - 2, ny, bcy);
        const int d = owner_dir1(mx, TB, ti, ntx, bcx) + 3 * owner_dir1(my, TB, tj, nty, bcyo);
        sc[lx + ly * W2] = (nb.par[0][d] ? b1 : b0)[mx + (int64_t)my * nx] - nb.shift[0][d];
    }
    __syncthreads();
    // 2. mu on the tile and a 1-cell ring (stored ring values of sleeping owners)
    int flag = 0;
    for (int k = tid; k < (w + 2) * (h + 2); k += NT) {
        const int lx = k % (w + 2), ly = k / (w + 2);
        const bool own = lx >= 1 && lx <= w && ly >= 1 && ly <= h;
        const int mx = map_index(x0 + lx - 1, nx, bcx), my = map_index(y0 + ly - 1, ny, bcy);
        const int d = own ? 4 : owner_dir1(mx, TB, ti, ntx, bcx) + 3 * owner_dir1(my, TB, tj, nty, bcyo);
        double mu;
        if (own || nb.act[0][d]) {
            const double lcv = NINE ? lds_lap9<W2>(sc, lx + 1, ly + 1, inv) : lds_lap5<W2>(sc, lx + 1, ly + 1, inv);
            const double cv = sc[(lx + 1) + (ly + 1) * W2];
            mu = (cv * cv * cv - cv) - kappa * lcv;
            if (own) flag |= fabs(lcv) > eps;
        } else {
            mu = TileRing<TB>::stale(mu_store, nb.tile[d], mx, my, nx, ny);
        }
        smu[lx + ly * W1] = mu;
    }
    __syncthreads();
    // 3. c' on the tile
    double part = 0.0;
    {
        const double dt = I.dt, mob = I.mobility, rg = I.react_gamma, cbar = I.cbar;
        for (int k = tid; k < w * h; k += NT) {
            const int lx = k % w, ly = k / w;
            const double lm = NINE ? lds_lap9<W1>(smu, lx + 1, ly + 1, inv) : lds_lap5<W1>(smu, lx + 1, ly + 1, inv);
            const double cv = sc[(lx + 2) + (ly + 2) * W2];
            const double change = dt * (mob * lm - rg * (cv - cbar));
            nxt[(x0 + lx) + (int64_t)(y0 + ly) * nx] = cv + change;
            part += change;
        }
    }
    // 4. detector bit; the ring for the neighbours when the tile did not fire
    const int any = regular ? 0 : block_any<NT>(flag, sflag);
    if (WORKLIST && !regular && tid == 0) I.nflag[0][t] = (uint8_t)any;
    if (!regular && !any) {
        for (int k = tid; k < TileRing<TB>::size(w, h); k += NT) {
            int lx, ly;
            TileRing<TB>::cell(k, &lx, &ly, w, h);
            mu_store[(int64_t)t * TileRing<TB>::N + k] = smu[(lx + 1) + (ly + 1) * W1];
        }
    }
    if (cons) {
        const double tot = block_sum<NT>(part, sred);
        if (tid == 0) {
            if (WORKLIST) {
                I.tpart[t] = tot;
                I.s_tile[(pself ^ 1) * I.nt + t] = I.scal[8];
            } else {
                nb.tot = tot;
            }
        }
    }
    __syncthreads();
    return any;
}

// Staging of a full tile for ch_tile: the thread's strip of c into registers
// (cout) and LDS, the 2-cell halo into LDS, from the buffers the neighbour
// table names, less the conserved shifts (the arithmetic of ch_tile's own
// staging). Every load is issued before the first wait: a halo slot past the
// region's count loads halo cell 0 again and stores nothing. (ch_tile's loop
// had waited for each load in turn, and a branch on the stencil between the
// strip's first load and the rest split them too.)
template <int TB, int NT>
__device__ __forceinline__ void ch_stage(CInst& I, int t, const NbTable& nb, double* sc, double* cout) {
    constexpr int W2 = TB + 4;
    constexpr int R = TB * TB / NT;
    constexpr int NH2 = W2 * W2 - TB * TB;
    constexpr int HK = (NH2 + NT - 1) / NT;
    const int tid = threadIdx.x, tx = tid % TB, ty = tid / TB;
    const int nx = I.nx, ny = I.ny, ntx = I.ntx, bcx = I.bc_x, bcy = I.bc_y;
    const int x0 = (t % ntx) * TB, y0 = (t / ntx) * TB;
    const gdouble* __restrict__ b0 = I.buf[0][0];
    const gdouble* __restrict__ b1 = I.buf[1][0];
    const gdouble* __restrict__ cur = nb.par[0][4] ? b1 : b0;
    const int gx = x0 + tx;
    double c[R], hv[HK], hs[HK];
    int hl[HK];
#pragma unroll
    for (int r = 0; r < R; r++) c[r] = cur[gx + (int64_t)(y0 + ty * R + r) * nx];
#pragma unroll
    for (int q = 0; q < HK; q++) {
        const int k0 = tid + q * NT;
        const int k = k0 < NH2 ? k0 : 0;
        int lx, ly;
        if (k < 2 * W2) { ly = k / W2; lx = k % W2; }                      // two rows below
        else if (k < 4 * W2) { ly = TB + 2 + (k - 2 * W2) / W2; lx = (k - 2 * W2) % W2; }  // two rows above
        else { const int q4 = k - 4 * W2; ly = 2 + q4 / 4; const int c4 = q4 % 4; lx = c4 < 2 ? c4 : TB + c4; }
        const int d = nb_dir1<TB, 2>(lx) + 3 * nb_dir1<TB, 2>(ly);
        const int mx = map_index(x0 + lx - 2, nx, bcx), my = map_index(y0 + ly - 2, ny, bcy);
        hv[q] = (nb.par[0][d] ? b1 : b0)[mx + (int64_t)my * nx];
        hs[q] = nb.shift[0][d];
        hl[q] = k0 < NH2 ? lx + ly * W2 : -1;
    }
    const double sh_self = nb.shift[0][4];
#pragma unroll
    for (int r = 0; r < R; r++) {
        cout[r] = c[r] - sh_self;
        sc[(tx + 2) + (ty * R + r + 2) * W2] = cout[r];
    }
#pragma unroll
    for (int q = 0; q < HK; q++)
        if (hl[q] >= 0) sc[hl[q]] = hv[q] - hs[q];
}

// Whether tile t of I is a partial last tile (only in the kernels' PART instances).
template <int TB, bool PART>
__device__ __forceinline__ bool ch_partial(CInst& I, int t) {
    return PART && !I.full_tiles && ((t % I.ntx) * TB + TB > I.nx || (t / I.ntx) * TB + TB > I.ny);
}

// One tile's step. WORKLIST: fill the neighbour table from the record and store
// the detector byte here; otherwise the caller has filled the table (then a
// barrier) and stores the returned detector bit any(|lap c| > eps) (0 in regular
// mode, uniform over the workgroup) itself.
// PART: the kernel serves instances with partial tiles (a separate instance of
// the kernels, so the partial path's registers do not cost full-tile grids).
// STAGED: the caller has filled the table (also in WORKLIST mode) and staged c
// with ch_stage; cin holds the thread's strip.
template <int TB, int NT, bool NINE, bool WORKLIST, bool PART = false, bool STAGED = false>
__device__ __forceinline__ int ch_tile(CInst& I, int t, uint64_t rec, NbTable& nb, double* sc, double* smu,
                                       double* sred, int* sflag, const double* cin = nullptr) {
    constexpr int W2 = TB + 4;
    constexpr int W1 = TB + 2;
    constexpr int R = TB * TB / NT;
    constexpr int NH2 = W2 * W2 - TB * TB;  // halo cells of the c region
    constexpr int NH1 = W1 * W1 - TB * TB;  // ring cells of the mu region
    const int tid = threadIdx.x;
    const int tx = tid % TB, ty = tid / TB;
    const int nx = I.nx, ny = I.ny, ntx = I.ntx, bcx = I.bc_x, bcy = I.bc_y;
    const int x0 = (t % ntx) * TB, y0 = (t / ntx) * TB;
    if (!STAGED && PART && !I.full_tiles && (x0 + TB > nx || y0 + TB > ny))
        return ch_tile_partial<TB, NT, NINE, WORKLIST>(I, t, rec, nb, sc, smu, sred, sflag);
    const bool regular = I.mode == 0;
    const bool cons = I.mode == 2;
    const gdouble* __restrict__ b0 = I.buf[0][0];
    const gdouble* __restrict__ b1 = I.buf[1][0];
    gdouble* __restrict__ mu_store = I.aux[0];
    const double kappa = I.kappa, eps = I.thresh_a;
    const double inv = NINE ? I.inv6dx2 : I.invdx2;
    if (WORKLIST && !STAGED) {
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
    if (STAGED) {
#pragma unroll
        for (int r = 0; r < R; r++) c[r] = cin[r];
    } else {
#pragma unroll
        for (int r = 0; r < R; r++) {
            c[r] = cur[gx + (int64_t)(y0 + ty * R + r) * nx] - sh_self;
            sc[(tx + 2) + (ty * R + r + 2) * W2] = c[r];
        }
    }
    if (!STAGED)
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
        }
    }
    // the detector's per-wave votes ride on the barrier after the mu ring (no
    // barriers of its own); sflag is next written after this tile's last barrier
    if (!regular) {
        const unsigned long long vote = __ballot(flag);
        if ((tid & 63) == 0) sflag[tid >> 6] = vote != 0ull;
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
            mu = TileRing<TB>::stale(mu_store, nb.tile[d], mx, my, nx, ny);
        }
        smu[lx + ly * W1] = mu;
    }
    __syncthreads();
    int any = 0;
    if (!regular) {
#pragma unroll
        for (int w = 0; w < NT / 64; w++) any |= sflag[w];
    }

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

    // 4. detector bit; the edge-ring mu for the neighbours, only when the tile
    //    did not fire: a tile that fired is awake next step (the dilation keeps
    //    the tile itself), and a ring is read only while its owner sleeps, so
    //    the ring a sleeping owner presents is the one of the last step it ran
    if (WORKLIST && !regular && tid == 0) I.nflag[0][t] = (uint8_t)any;
    if (!regular && !any) {
        for (int k = tid; k < TileRing<TB>::N; k += NT) {
            int lx, ly;
            TileRing<TB>::cell(k, &lx, &ly);
            mu_store[(int64_t)t * TileRing<TB>::N + k] = smu[(lx + 1) + (ly + 1) * W1];
        }
    }
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
    if (WORKLIST) __syncthreads();  // LDS reuse by the next item; a single-launch tile is the last
    return any;
}

template <int TB, int NT, bool PART>
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
        if (ch_partial<TB, PART>(I, t)) {
            if (I.stencil != 5) ch_tile<TB, NT, true, true, PART>(I, t, rec, nb, sc, smu, sred, sflag);
            else ch_tile<TB, NT, false, true, PART>(I, t, rec, nb, sc, smu, sred, sflag);
            continue;
        }
        nb_fill<TB>(nb, I, t, 0, 0, I.mode == 0, 1, I.mode == 2 ? 1 : 0, rec);
        __syncthreads();
        double c[TB * TB / NT];
        ch_stage<TB, NT>(I, t, nb, sc, c);
        if (I.stencil != 5) ch_tile<TB, NT, true, true, PART, true>(I, t, rec, nb, sc, smu, sred, sflag, c);
        else ch_tile<TB, NT, false, true, PART, true>(I, t, rec, nb, sc, smu, sred, sflag, c);
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
template <int TB, int NT, bool CONS, bool PART>
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
    dense_prefetch(I, 1);
    const int prev = s & 1, next = prev ^ 1;
    const bool regular = I.mode == 0;
    constexpr bool cons = CONS;  // conserved DBA: one instance in one segment (the host checks)
    const int runs = load_early(I.druns + t);  // read early, off the tail (every lane: no branch)
    // The previous step's change sums (tpart + (1 + parity) nt, 0 for tiles that
    // did not run) in the finalize's order: 16 words of 64 tiles, each by the
    // shuffle tree of block_sum, then the words in order. Their loads go out
    // before the neighbour table's, so the two round trips overlap.
    constexpr int NWV = NT / 64, PER = (kFinNT / 64 + NWV - 1) / NWV;
    double v[PER];
    if (cons && !first) {
        const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
        const gdouble* tp = I.tpart + (int64_t)(1 + prev) * I.nt;
#pragma unroll
        for (int k = 0; k < PER; k++) {  // all of this wave's words in flight at once
            const int w = wid + k * NWV, tt = w * 64 + lane;
            v[k] = (w < kFinNT / 64 && tt < I.nt) ? tp[tt] : 0.0;
        }
    }
    dense_nb_fill(nb, I, t, first, prev, regular, 1, cons);
    if (cons && !first) {
        const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
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
    if (ran) {
        if (ch_partial<TB, PART>(I, t)) {
            any = I.stencil != 5 ? ch_tile<TB, NT, true, false, PART>(I, t, 0, nb, sc, smu, sred, sflag)
                                 : ch_tile<TB, NT, false, false, PART>(I, t, 0, nb, sc, smu, sred, sflag);
        } else {
            double c[TB * TB / NT];
            ch_stage<TB, NT>(I, t, nb, sc, c);
            any = I.stencil != 5 ? ch_tile<TB, NT, true, false, PART, true>(I, t, 0, nb, sc, smu, sred, sflag, c)
                                 : ch_tile<TB, NT, false, false, PART, true>(I, t, 0, nb, sc, smu, sred, sflag, c);
        }
    }
    if (threadIdx.x == 0) {
        I.dF[next][t] = (uint8_t)any;
        I.dP[next][t] = (uint8_t)(pself ^ ran);
        if (ran) {
            I.druns[t] = runs + 1;
            if (I.hist_cap > 0) {
                const int64_t step = I.ctr[0] + s;
                if (step < I.hist_cap) atomicAdd((int*)&I.hist[2 * step], 1);
            }
        }
        if (cons) {
            I.tpart[(int64_t)(1 + next) * I.nt + t] = ran ? nb.tot : 0.0;
            if (ran) I.s_tile[(pself ^ 1) * I.nt + t] = nb.S;
        }
    }
}

}  // namespace dba

namespace dba {

// ---- Column-march single-launch step for large CH grids ----
//
// One workgroup walks a strip of K tiles up one tile column (32 x 32 tiles, 256
// threads, periodic both ways, one instance, not conserved). The c rows of the
// strip live in a 64-row LDS ring: tile j reads ring rows 32 j .. 32 j + 35
// (its 32 rows and 2 halo rows below and above), and the next tile reuses the
// 4 rows the two share, so every row of the strip is read from memory once.
// While tile j computes, the 32 new rows of tile j + 1 (with their 2-cell x
// halo) are already loading into registers; the barriers between the phases
// order LDS only, so those loads stay in flight across them. Per tile the
// arithmetic, the operation order, the detector, the stale mu of sleeping
// neighbours and the bytes written are those of ch_dense_kernel, so the two
// give identical results.

// Workgroup barrier that orders LDS only: outstanding global loads stay in flight.
__device__ __forceinline__ void lds_barrier() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

constexpr int kMarchRing = 64;  // LDS ring rows (a power of two >= 36 + 32 - 4)
constexpr int kMarchMinTiles = 4096;  // grids from 2048^2 (32 x 32 tiles) march

// 9- or 5-point Laplacian at window cell (x, y) of the ring (window row 0 at ring row base)
template <bool NINE>
__device__ __forceinline__ double ring_lap(const double* ring, int base, int x, int y, double inv) {
    constexpr int W2 = 36;
    const double* rm = ring + ((base + y - 1) & (kMarchRing - 1)) * W2;
    const double* r0 = ring + ((base + y) & (kMarchRing - 1)) * W2;
    const double* rp = ring + ((base + y + 1) & (kMarchRing - 1)) * W2;
    const double s1 = (r0[x + 1] + r0[x - 1]) + (rp[x] + rm[x]);
    if (NINE) {
        const double s2 = (rp[x + 1] + rp[x - 1]) + (rm[x + 1] + rm[x - 1]);
        return ((4.0 * s1 + s2) - 20.0 * r0[x]) * inv;
    }
    return (s1 - 4.0 * r0[x]) * inv;
}

template <int K, bool NINE>
__global__ __launch_bounds__(256) void ch_march_kernel(const Inst* __restrict__ insts, int first, int s) {
    constexpr int TB = 32, NT = 256, R = 4;
    constexpr int W2 = TB + 4, W1 = TB + 2;
    constexpr int NH1 = W1 * W1 - TB * TB;
    constexpr int NC = 3 * (K + 2);  // columns (left, own, right) x tile rows ty0 - 1 .. ty0 + K
    static_assert(NC <= 128 && K <= NT - 128, "table threads below 128, run-count threads from 128");
    constexpr int MASK = kMarchRing - 1;
    __shared__ double ring[kMarchRing * W2];
    __shared__ double smu[W1 * W1];
    __shared__ int s_tile[NC], s_par[NC], s_act[NC];
    __shared__ int s_runs[K];
    __shared__ int sflag[NT / 64 + 1];
    CInst& I = ((CInst*)insts)[0];
    dense_prefetch(I, 1);
    const int tid = threadIdx.x, tx = tid % TB, ty = tid / TB;
    const int nx = I.nx, ny = I.ny, ntx = I.ntx, nty = I.nty;
    const int prev = s & 1, next = prev ^ 1;
    const bool regular = I.mode == 0;
    const gdouble* __restrict__ b0 = I.buf[0][0];
    const gdouble* __restrict__ b1 = I.buf[1][0];
    gdouble* __restrict__ mu_store = I.aux[0];
    const double kappa = I.kappa, eps = I.thresh_a;
    const double inv = NINE ? I.inv6dx2 : I.invdx2;
    const double dt = I.dt, mob = I.mobility, rg = I.react_gamma, cbar = I.cbar;
    const int item = xcd_deal(blockIdx.x, gridDim.x, 0, gridDim.x);
    const int sx = item % ntx, ty0 = (item / ntx) * K;
    const int x0 = sx * TB;

    // 0. owner, buffer parity and awake bit of the tiles around the strip
    if (tid < NC) {
        const int c = tid % 3, r = tid / 3;
        const int oi = sx + c - 1 < 0 ? ntx - 1 : (sx + c - 1 >= ntx ? 0 : sx + c - 1);
        const int oj0 = ty0 + r - 1;
        const int oj = oj0 < 0 ? oj0 + nty : (oj0 >= nty ? oj0 - nty : oj0);
        const int o = oi + oj * ntx;
        int par, act = 1;
        if (first) {
            par = bit_of(I.par[0], o);
            if (!regular) act = bit_of(I.a_cur[0], o);
        } else {
            const gu8 *F, *P;
            dense_bytes(I, prev, &F, &P);
            par = P[o];
            if (!regular) act = dilated2d(F, oi, oj, ntx, nb2_of(oi, oj, ntx, nty, 0, 0), 0);
        }
        s_tile[tid] = o;
        s_par[tid] = par;
        s_act[tid] = act;
    } else if (tid >= 128 && tid < 128 + K) {
        // the strip's per-tile update counts, read now so no tile waits on them
        // (a wave of their own, so they load alongside the table)
        s_runs[tid - 128] = I.druns[sx + (ty0 + tid - 128) * ntx];
    }
    __syncthreads();

    // Rows [L, L + 32) of the strip's window (global row ty0 * 32 - 2 + L): the
    // thread's 4 rows of its column, and (threads 0..127) one x-halo cell.
    const int hq = tid >> 2, hc = tid & 3, hlx = hc < 2 ? hc : TB + hc;  // halo row, LDS column
    const int hmx = map_index(x0 + hlx - 2, nx, 0);
    auto row_of = [&](int L) { return map_index(ty0 * TB - 2 + L, ny, 0); };
    auto trow = [&](int L) { return (L + TB - 2) / TB; };  // table row of window row L
    double pm[R], ph = 0.0;
    auto fetch = [&](int L) {
#pragma unroll
        for (int r = 0; r < R; r++) {
            const int Lr = L + ty * R + r;
            const gdouble* src = s_par[1 + 3 * trow(Lr)] ? b1 : b0;
            pm[r] = src[x0 + tx + (int64_t)row_of(Lr) * nx];
        }
        if (tid < 4 * TB) {
            const int Lr = L + hq;
            const gdouble* src = s_par[(hc < 2 ? 0 : 2) + 3 * trow(Lr)] ? b1 : b0;
            ph = src[hmx + (int64_t)row_of(Lr) * nx];
        }
    };
    auto deposit = [&](int L) {
#pragma unroll
        for (int r = 0; r < R; r++) ring[((L + ty * R + r) & MASK) * W2 + tx + 2] = pm[r];
        if (tid < 4 * TB) ring[((L + hq) & MASK) * W2 + hlx] = ph;
    };
    // the 4 window rows below the first block (2 halo rows, the first tile's rows 0-1)
    if (tid < 4 * W2) {
        const int L = tid / W2, lx = tid % W2;
        const int c = lx < 2 ? 0 : (lx < TB + 2 ? 1 : 2);
        const gdouble* src = s_par[c + 3 * trow(L)] ? b1 : b0;
        ring[(L & MASK) * W2 + lx] = src[map_index(x0 + lx - 2, nx, 0) + (int64_t)row_of(L) * nx];
    }
    // block j covers window rows [32 j + 4, 32 j + 36): tile j's rows 2-31 and
    // tile j + 1's rows 0-1; it is needed when tile j or tile j + 1 runs
    auto needed = [&](int j) { return (s_act[1 + 3 * (j + 1)] | s_act[1 + 3 * (j + 2)]) != 0; };
    if (needed(0)) {
        fetch(4);
        deposit(4);
    }
    __syncthreads();

#pragma unroll 1
    for (int j = 0; j < K; j++) {
        const bool pre = j + 1 < K && needed(j + 1);
        if (pre) fetch(32 * (j + 1) + 4);  // in flight while tile j computes
        const int tr = j + 1;               // table row of tile j
        const int t = sx + (ty0 + j) * ntx;
        const int pself = s_par[1 + 3 * tr];
        const int ran = s_act[1 + 3 * tr];
        const int base = 32 * j;            // ring row of window row 0 of tile j
        int any = 0;
        if (ran) {
            gdouble* __restrict__ nxt = pself ? I.buf[0][0] : I.buf[1][0];
            // mu on the strip (+ detector), then on the 1-cell ring
            double c[R];
            int flag = 0;
            {
                double Lc[R + 2], Cc[R + 2], Rc[R + 2];
#pragma unroll
                for (int r = 0; r < R + 2; r++) {
                    const double* row = ring + ((base + ty * R + 1 + r) & MASK) * W2;
                    Lc[r] = row[tx + 1];
                    Cc[r] = row[tx + 2];
                    Rc[r] = row[tx + 3];
                }
#pragma unroll
                for (int r = 0; r < R; r++) {
                    const double s1 = (Rc[r + 1] + Lc[r + 1]) + (Cc[r + 2] + Cc[r]);
                    double lc;
                    if (NINE) {
                        const double s2 = (Rc[r + 2] + Lc[r + 2]) + (Rc[r] + Lc[r]);
                        lc = ((4.0 * s1 + s2) - 20.0 * Cc[r + 1]) * inv;
                    } else {
                        lc = (s1 - 4.0 * Cc[r + 1]) * inv;
                    }
                    const double cv = Cc[r + 1];
                    c[r] = cv;
                    smu[(tx + 1) + (ty * R + r + 1) * W1] = (cv * cv * cv - cv) - kappa * lc;
                    flag |= fabs(lc) > eps;
                }
            }
            for (int k = tid; k < NH1; k += NT) {
                int lx, ly;
                if (k < W1) { ly = 0; lx = k; }
                else if (k < 2 * W1) { ly = TB + 1; lx = k - W1; }
                else { const int q = k - 2 * W1; ly = 1 + q / 2; lx = (q % 2) ? TB + 1 : 0; }
                const int dxi = nb_dir1<TB, 1>(lx), dyi = nb_dir1<TB, 1>(ly);
                const int e = dxi + 3 * (tr + dyi - 1);
                double mu;
                if (s_act[e]) {
                    const double lcv = ring_lap<NINE>(ring, base, lx + 1, ly + 1, inv);
                    const double cv = ring[((base + ly + 1) & MASK) * W2 + lx + 1];
                    mu = (cv * cv * cv - cv) - kappa * lcv;
                } else {
                    const int mx = map_index(x0 + lx - 1, nx, 0);
                    const int my = map_index((ty0 + j) * TB + ly - 1, ny, 0);
                    mu = TileRing<TB>::stale(mu_store, s_tile[e], mx, my, nx, ny);
                }
                smu[lx + ly * W1] = mu;
            }
            if (!regular) {  // detector votes ride on this barrier (see ch_tile)
                const unsigned long long vote = __ballot(flag);
                if ((tid & 63) == 0) sflag[tid >> 6] = vote != 0ull;
            }
            lds_barrier();
            int anyv = 0;
            if (!regular) {
#pragma unroll
                for (int w = 0; w < NT / 64; w++) anyv |= sflag[w];
            }
            // c' on the strip
            {
                double Lm[R + 2], Cm[R + 2], Rm[R + 2];
#pragma unroll
                for (int r = 0; r < R + 2; r++) {
                    const int row = (ty * R + r) * W1;
                    Lm[r] = smu[row + tx];
                    Cm[r] = smu[row + tx + 1];
                    Rm[r] = smu[row + tx + 2];
                }
#pragma unroll
                for (int r = 0; r < R; r++) {
                    const double s1 = (Rm[r + 1] + Lm[r + 1]) + (Cm[r + 2] + Cm[r]);
                    double lm;
                    if (NINE) {
                        const double s2 = (Rm[r + 2] + Lm[r + 2]) + (Rm[r] + Lm[r]);
                        lm = ((4.0 * s1 + s2) - 20.0 * Cm[r + 1]) * inv;
                    } else {
                        lm = (s1 - 4.0 * Cm[r + 1]) * inv;
                    }
                    const double change = dt * (mob * lm - rg * (c[r] - cbar));
                    nxt[x0 + tx + (int64_t)((ty0 + j) * TB + ty * R + r) * nx] = c[r] + change;
                }
            }
            any = anyv;
            // tile j read its window (ring rows 32 j .. 32 j + 35) only before the
            // mu barrier, so the next tile's rows, whose last 4 wrap onto this
            // window's first 4, can be written now
            if (pre) deposit(32 * (j + 1) + 4);
            if (!regular && !any) {
                for (int k = tid; k < TileRing<TB>::N; k += NT) {
                    int lx, ly;
                    TileRing<TB>::cell(k, &lx, &ly);
                    mu_store[(int64_t)t * TileRing<TB>::N + k] = smu[(lx + 1) + (ly + 1) * W1];
                }
            }
        } else if (pre) {
            deposit(32 * (j + 1) + 4);  // a sleeping tile reads no ring rows
        }
        if (tid == 0) {
            I.dF[next][t] = (uint8_t)any;
            I.dP[next][t] = (uint8_t)(pself ^ ran);
            if (ran) {
                I.druns[t] = s_runs[j] + 1;
                if (I.hist_cap > 0) {
                    const int64_t step = I.ctr[0] + s;
                    if (step < I.hist_cap) atomicAdd((int*)&I.hist[2 * step], 1);
                }
            }
        }
        // the next tile's rows were deposited after the mu barrier (below the
        // c' pass) or at once by a sleeping tile; this barrier makes them
        // visible and ends every read of mu
        lds_barrier();
    }
}

}  // namespace dba
