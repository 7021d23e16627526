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
#include "k_finalize.h"

namespace dba {

__device__ __forceinline__ void k93_aniso(CInst& I, double gx, double gy, double* eps, double* a) {
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

// LDS of one K93 tile.
template <int TB>
struct K93Lds {
    static constexpr int W2 = TB + 4;
    static constexpr int W1 = TB + 2;
    double sp[W2 * W2];
    double sT[W1 * W1];
    double sFx[W1 * W1];
    double sFy[W1 * W1];
    double sE2[TB * TB];
};

// One tile's step (two sets: p, T). WORKLIST: fill the neighbour table from the
// record and store the detector bytes here; otherwise the caller has filled the
// table and stores the returned bits (p) | (T) << 1 (uniform over the
// workgroup, 0 in regular mode; bit s is meaningful only for a set that ran),
// and, the workgroup's only tile, it skips the closing barriers.
// PARTIAL: a partial last tile (w x h cells) when the grid is not a multiple of
// the tile; the owner of each halo cell then comes from its mapped position.
// STAGED: the caller has already written sp / sT (k93_stage_select).
template <int TB, int NT, bool WORKLIST, bool PARTIAL = false, bool STAGED = false>
__device__ __forceinline__ int k93_tile(CInst& I, int t, uint64_t rec, NbTable& nb, K93Lds<TB>& L,
                                        int* sflag) {
    constexpr int W2 = TB + 4;
    constexpr int W1 = TB + 2;
    double* sp = L.sp;
    double* sT = L.sT;
    double* sFx = L.sFx;
    double* sFy = L.sFy;
    double* sE2 = L.sE2;
    const int tid = threadIdx.x;
    const int nx = I.nx, ny = I.ny, ntx = I.ntx, nty = I.nty, bcx = I.bc_x, bcy = I.bc_y;
    const int ti = t % ntx, tj = t / ntx;
    const int x0 = ti * TB, y0 = tj * TB;
    // tile extent (TB x TB unless partial) and the owner direction of a mapped cell
    const int w = PARTIAL ? tile_extent(ti, TB, nx) : TB, h = PARTIAL ? tile_extent(tj, TB, ny) : TB;
    const int bcyo = bcy == 0 ? 0 : 1;
    auto odir = [&](int mx, int my) {
        return owner_dir1(mx, TB, ti, ntx, bcx) + 3 * owner_dir1(my, TB, tj, nty, bcyo);
    };
    const bool regular = I.mode == 0;
    const gdouble* __restrict__ p0 = I.buf[0][0];
    const gdouble* __restrict__ p1 = I.buf[1][0];
    const gdouble* __restrict__ T0 = I.buf[0][1];
    const gdouble* __restrict__ T1 = I.buf[1][1];
    if (WORKLIST) {
        nb_fill<TB>(nb, I, t, x0, y0, regular, 2, 0, rec);
        __syncthreads();
    }
    const int ap = nb.act[0][4], at = nb.act[1][4];
    const int pp0 = nb.par[0][4], pp1 = nb.par[1][4];
    gdouble* __restrict__ pnew = pp0 ? I.buf[0][0] : I.buf[1][0];
    gdouble* __restrict__ Tnew = pp1 ? I.buf[0][1] : I.buf[1][1];

    // stage phi (2-cell halo) when p-awake, T (1-cell halo) when T-awake,
    // and phi / T of the tile itself always (the T update needs phi' - phi).
    // A full tile issues every load of the thread's cells before the first wait
    // (a cell not needed, or past the region, loads a valid cell and is not
    // stored); a partial tile takes the generic loops below.
    if (!STAGED && !PARTIAL) {
        constexpr int KP = (W2 * W2 + NT - 1) / NT, KT = (W1 * W1 + NT - 1) / NT;
        double vp[KP], vT[KT];
#pragma unroll
        for (int q = 0; q < KP; q++) {
            const int k0 = tid + q * NT, k = k0 < W2 * W2 ? k0 : 0;
            const int lx = k % W2, ly = k / W2;
            const int mx = map_index(x0 + lx - 2, nx, bcx), my = map_index(y0 + ly - 2, ny, bcy);
            const int d = nb_dir1<TB, 2>(lx) + 3 * nb_dir1<TB, 2>(ly);
            vp[q] = (nb.par[0][d] ? p1 : p0)[mx + (int64_t)my * nx];
        }
#pragma unroll
        for (int q = 0; q < KT; q++) {
            const int k0 = tid + q * NT, k = k0 < W1 * W1 ? k0 : 0;
            const int lx = k % W1, ly = k / W1;
            const int mx = map_index(x0 + lx - 1, nx, bcx), my = map_index(y0 + ly - 1, ny, bcy);
            const int d = nb_dir1<TB, 1>(lx) + 3 * nb_dir1<TB, 1>(ly);
            vT[q] = (nb.par[1][d] ? T1 : T0)[mx + (int64_t)my * nx];
        }
#pragma unroll
        for (int q = 0; q < KP; q++) {
            const int k = tid + q * NT;
            const int lx = k % W2, ly = k / W2;
            const bool own = lx >= 2 && lx < TB + 2 && ly >= 2 && ly < TB + 2;
            if (k < W2 * W2 && (own || ap)) sp[k] = vp[q];
        }
#pragma unroll
        for (int q = 0; q < KT; q++) {
            const int k = tid + q * NT;
            const int lx = k % W1, ly = k / W1;
            const bool own = lx >= 1 && lx <= TB && ly >= 1 && ly <= TB;
            if (k < W1 * W1 && (own || at)) sT[k] = vT[q];
        }
    }
    if (!STAGED && PARTIAL)
    for (int k = tid; k < (w + 4) * (h + 4); k += NT) {
        const int lx = k % (w + 4), ly = k / (w + 4);
        const bool own = lx >= 2 && lx < w + 2 && ly >= 2 && ly < h + 2;
        if (!own && !ap) continue;
        const int mx = map_index(x0 + lx - 2, nx, bcx), my = map_index(y0 + ly - 2, ny, bcy);
        const int d = PARTIAL ? (own ? 4 : odir(mx, my)) : nb_dir1<TB, 2>(lx) + 3 * nb_dir1<TB, 2>(ly);
        sp[lx + ly * W2] = (nb.par[0][d] ? p1 : p0)[mx + (int64_t)my * nx];
    }
    if (!STAGED && PARTIAL)
    for (int k = tid; k < (w + 2) * (h + 2); k += NT) {
        const int lx = k % (w + 2), ly = k / (w + 2);
        const bool own = lx >= 1 && lx <= w && ly >= 1 && ly <= h;
        if (!own && !at) continue;
        const int mx = map_index(x0 + lx - 1, nx, bcx), my = map_index(y0 + ly - 1, ny, bcy);
        const int d = PARTIAL ? (own ? 4 : odir(mx, my)) : nb_dir1<TB, 1>(lx) + 3 * nb_dir1<TB, 1>(ly);
        sT[lx + ly * W1] = (nb.par[1][d] ? T1 : T0)[mx + (int64_t)my * nx];
    }
    __syncthreads();

    // eps eps' grad phi on the tile + ring, eps^2 on the tile
    if (ap) {
        const double inv2dx = I.inv2dx;
        for (int k = tid; k < (w + 2) * (h + 2); k += NT) {
            const int lx = k % (w + 2), ly = k / (w + 2);
            const int cx = lx + 1, cy = ly + 1;
            const double gx = (sp[(cx + 1) + cy * W2] - sp[(cx - 1) + cy * W2]) * inv2dx;
            const double gy = (sp[cx + (cy + 1) * W2] - sp[cx + (cy - 1) * W2]) * inv2dx;
            double e, a;
            k93_aniso(I, gx, gy, &e, &a);
            sFx[lx + ly * W1] = a * gx;
            sFy[lx + ly * W1] = a * gy;
            if (lx >= 1 && lx <= w && ly >= 1 && ly <= h) sE2[(lx - 1) + (ly - 1) * TB] = e * e;
        }
    }
    __syncthreads();

    int fp = 0, fT = 0;
    const double inv2dx = I.inv2dx, invdx2 = I.invdx2, alpha_pi = I.alpha_pi, gam = I.gamma;
    const double t_eq = I.t_eq, dt_tau = I.dt_tau, dt = I.dt, lk = I.latent_k;
    const double mp = I.thresh_a, mT = I.thresh_b;
    for (int k = tid; k < w * h; k += NT) {
        const int lx = k % w, ly = k / w;
        const int64_t gi = (x0 + lx) + (int64_t)(y0 + ly) * nx;
        const double ph = sp[(lx + 2) + (ly + 2) * W2];
        const double Tv = sT[(lx + 1) + (ly + 1) * W1];
        double pn = ph;
        if (ap) {
            const int cx = lx + 2, cy = ly + 2, fx = lx + 1, fy = ly + 1;
            const double cross = ((sFx[fx + (fy + 1) * W1] - sFx[fx + (fy - 1) * W1]) -
                                  (sFy[(fx + 1) + fy * W1] - sFy[(fx - 1) + fy * W1])) * inv2dx;
            const double lap = lds_lap5<W2>(sp, cx, cy, invdx2);
            const double m = alpha_pi * atan(gam * (t_eq - Tv));
            const double rhs = (cross + sE2[lx + ly * TB] * lap) + (ph * (1.0 - ph)) * ((ph - 0.5) + m);
            pn = ph + dt_tau * rhs;
            pnew[gi] = pn;
            fp |= (pn >= mp && pn <= (1.0 - mp));
        }
        if (at) {
            const double lapT = lds_lap5<W1>(sT, lx + 1, ly + 1, invdx2);
            const double t1 = Tv + dt * lapT;
            const double Tn = ap ? t1 + lk * (pn - ph) : t1;
            Tnew[gi] = Tn;
            fT |= fabs(Tn - Tv) >= mT;
        }
    }
    int bits = 0;
    if (!regular) {
        bits = block_any2<NT, WORKLIST>(fp, fT, sflag);
        const int anyp = bits & 1, anyT = bits >> 1;
        if (WORKLIST && tid == 0) {
            if (ap) I.nflag[0][t] = (uint8_t)anyp;
            if (at) I.nflag[1][t] = (uint8_t)anyT;
        }
    }
    if (WORKLIST) __syncthreads();
    return bits;
}

// A partial last tile of an instance whose grid is not a multiple of the tile
// (only in the kernels' PART instances, which groups holding such an instance
// use, so the partial path's registers do not cost full-tile grids).
template <int TB, bool PART>
__device__ __forceinline__ bool k93_partial(CInst& I, int t) {
    return PART && !I.full_tiles && ((t % I.ntx) * TB + TB > I.nx || (t / I.ntx) * TB + TB > I.ny);
}

template <int TB, int NT, bool PART>
// 4 waves per SIMD: the staging's loads in flight (7 doubles per thread at 128
// threads) fit without spills (5 waves spilled 48 B per lane and ran slower)
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(4))) void k93_step_kernel(const WorkList wl) {
    __shared__ K93Lds<TB> L;
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
        if (k93_partial<TB, PART>(*Ip, t)) k93_tile<TB, NT, true, true>(*Ip, t, rec, nb, L, sflag);
        else k93_tile<TB, NT, true>(*Ip, t, rec, nb, L, sflag);
    }
}

// Speculative staging of a full tile (single-launch steps of small grids): the
// phi ring (2 cells) and T ring (1 cell) of both buffers are loaded into
// registers before the neighbour table is known, so the table's loads and these
// overlap; k93_stage_select then keeps, per cell, the buffer the owner's parity
// names. The extra bytes are the other buffer's copy of the rings of every
// tile, sleeping ones included, which is why only grids of up to 1024 tiles
// (all workgroups resident at once, the grid in the caches) take this path.
template <int TB, int NT>
struct K93Spec {
    static constexpr int W2 = TB + 4, W1 = TB + 2;
    static constexpr int KP = (W2 * W2 + NT - 1) / NT, KT = (W1 * W1 + NT - 1) / NT;
    double p[2][KP], T[2][KT];
};

template <int TB, int NT>
__device__ __forceinline__ void k93_stage_load(CInst& I, int t, K93Spec<TB, NT>& r) {
    using S = K93Spec<TB, NT>;
    const int ntx = I.ntx, nx = I.nx, ny = I.ny, bcx = I.bc_x, bcy = I.bc_y;
    const int ti = t % ntx, tj = t / ntx, x0 = ti * TB, y0 = tj * TB;
#pragma unroll
    for (int q = 0; q < S::KP; q++) {
        const int k = threadIdx.x + q * NT;
        if (k < S::W2 * S::W2) {
            const int lx = k % S::W2, ly = k / S::W2;
            const int64_t gi = map_index(x0 + lx - 2, nx, bcx) + (int64_t)map_index(y0 + ly - 2, ny, bcy) * nx;
            r.p[0][q] = I.buf[0][0][gi];
            r.p[1][q] = I.buf[1][0][gi];
        }
    }
#pragma unroll
    for (int q = 0; q < S::KT; q++) {
        const int k = threadIdx.x + q * NT;
        if (k < S::W1 * S::W1) {
            const int lx = k % S::W1, ly = k / S::W1;
            const int64_t gi = map_index(x0 + lx - 1, nx, bcx) + (int64_t)map_index(y0 + ly - 1, ny, bcy) * nx;
            r.T[0][q] = I.buf[0][1][gi];
            r.T[1][q] = I.buf[1][1][gi];
        }
    }
}

// The staging of k93_tile from the speculative registers: phi of the tile
// always and its ring when p-awake, T likewise when T-awake.
template <int TB, int NT>
__device__ __forceinline__ void k93_stage_select(const NbTable& nb, const K93Spec<TB, NT>& r, K93Lds<TB>& L) {
    using S = K93Spec<TB, NT>;
    const int ap = nb.act[0][4], at = nb.act[1][4];
#pragma unroll
    for (int q = 0; q < S::KP; q++) {
        const int k = threadIdx.x + q * NT;
        if (k < S::W2 * S::W2) {
            const int lx = k % S::W2, ly = k / S::W2;
            const bool own = lx >= 2 && lx < TB + 2 && ly >= 2 && ly < TB + 2;
            if (own || ap) L.sp[k] = nb.par[0][nb_dir1<TB, 2>(lx) + 3 * nb_dir1<TB, 2>(ly)] ? r.p[1][q] : r.p[0][q];
        }
    }
#pragma unroll
    for (int q = 0; q < S::KT; q++) {
        const int k = threadIdx.x + q * NT;
        if (k < S::W1 * S::W1) {
            const int lx = k % S::W1, ly = k / S::W1;
            const bool own = lx >= 1 && lx <= TB && ly >= 1 && ly <= TB;
            if (own || at) L.sT[k] = nb.par[1][nb_dir1<TB, 1>(lx) + 3 * nb_dir1<TB, 1>(ly)] ? r.T[1][q] : r.T[0][q];
        }
    }
}

// Single-launch step of a K93 group (see ch_dense_kernel): per
// set, awake bits and parities of the 3x3 owner tiles from the bitmaps (first
// step) or the previous step's bytes; the tile runs when either set is awake.
// SPEC: one instance of full tiles, staged speculatively (K93Spec). The tile's
// run counter is read at the start, off the path of the tile's last store.
template <int TB, int NT, bool PART, bool SPEC = false>
__global__ __launch_bounds__(NT) void k93_dense_kernel(const Inst* __restrict__ insts,
                                                       const int32_t* __restrict__ dmap, int first, int s) {
    __shared__ K93Lds<TB> L;
    __shared__ int sflag[NT / 64 + 1];
    __shared__ NbTable nb;
    int t;
    CInst& I = dense_item(insts, dmap, &t, false);
    dense_prefetch(I, 2);
    const int nt = I.nt;
    const int prev = s & 1, next = prev ^ 1;
    const bool regular = I.mode == 0;
    const int runs = load_early(I.druns + t);  // every lane: no branch to wait at
    K93Spec<TB, NT> spec;
    if (SPEC) k93_stage_load<TB, NT>(I, t, spec);
    dense_nb_fill(nb, I, t, first, prev, regular, 2, false);
    __syncthreads();
    const int ap = nb.act[0][4], at = nb.act[1][4];
    int bits = 0;
    if (ap | at) {
        if (SPEC) {
            k93_stage_select<TB, NT>(nb, spec, L);
            bits = k93_tile<TB, NT, false, false, true>(I, t, 0, nb, L, sflag);
        } else {
            bits = k93_partial<TB, PART>(I, t) ? k93_tile<TB, NT, false, true>(I, t, 0, nb, L, sflag)
                                               : k93_tile<TB, NT, false>(I, t, 0, nb, L, sflag);
        }
    }
    if (threadIdx.x == 0) {
        I.dF[next][t] = (uint8_t)(ap ? bits & 1 : 0);
        I.dF[next][nt + t] = (uint8_t)(at ? bits >> 1 : 0);
        I.dP[next][t] = (uint8_t)(nb.par[0][4] ^ ap);
        I.dP[next][nt + t] = (uint8_t)(nb.par[1][4] ^ at);
        if (ap | at) {
            I.druns[t] = runs + 1 + ap;  // field-set updates: T over A_p | A_T, p over A_p
            if (I.hist_cap > 0) {
                const int64_t step = I.ctr[0] + s;
                if (step < I.hist_cap) {
                    if (ap) atomicAdd((int*)&I.hist[2 * step], 1);
                    if (at) atomicAdd((int*)&I.hist[2 * step + 1], 1);
                }
            }
        }
    }
}

}  // namespace dba
