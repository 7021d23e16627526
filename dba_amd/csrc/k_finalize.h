// k_finalize.h - end-of-step activation bookkeeping and worklist compaction.
//
// Reference counterpart: the tail of every DBA step, Wakeup_Next (clears the
// flags of the blocks that ran, example_Dendritic_growth/kernel.cu:496-574)
// and Wakeup_Neighbor (sets each flagged block and its neighbours,
// kernel.cu:577-634; 3x3 in KH, example_Kelvin_Helmholtz/KH_DBA.cu:988-1044),
// plus the conserved mean shift (example_Reactive_ModelB/Conserved/kernel.cu:1152-1164).
//
// One workgroup per segment of <= 1024 tiles:
//   (a) counts the tiles that ran (history), flips their buffer parity, and
//       sums their conserved change partials;
//   (b) dilates the detector bytes of the step into the next awake set: each
//       lane evaluates one tile, a wave ballot gives a 64-tile bitmap word;
//   (c) compacts the segment's union awake set into an ascending list
//       (popcount prefix over the segment's words, then the lane's rank in
//       its word), i.e. the order copy_if over tile indices gives;
//   (d) the last workgroup to finish reduces the segment partials per
//       instance in segment order (history, counters, conserved shift, KH CFL
//       dt) and writes the exclusive prefix of the segment counts.
#pragma once
#include "common.h"

namespace dba {

__device__ __forceinline__ uint64_t full_word(int w, int nt) {
    int n = nt - w * 64;
    if (n >= 64) return ~0ull;
    if (n <= 0) return 0ull;
    return (1ull << n) - 1ull;
}

__device__ __forceinline__ int64_t tile_cells(const Inst& I, int t) {
    int ti = t % I.ntx, tj = (t / I.ntx) % I.nty, tk = t / (I.ntx * I.nty);
    int wx = min(I.tx, I.nx - ti * I.tx), wy = min(I.ty, I.ny - tj * I.ty), wz = min(I.tz, I.nz - tk * I.tz);
    return (int64_t)wx * wy * wz;
}

// Awake flag of tile t for the next step: OR of the detector bytes over the
// tile and its neighbours (kind 0: faces; kind 1: 3x3 in-plane + z faces),
// wrapping where the instance is periodic and clamping elsewhere.
__device__ __forceinline__ int dilated3d(const Inst& I, const uint8_t* N, int t, int kind) {
    const int ti = t % I.ntx, tj = (t / I.ntx) % I.nty, tk = t / (I.ntx * I.nty);
    int r = 0;
    for (int dk = -1; dk <= 1; dk++) {
        int kk = tk + dk;
        if (kk < 0 || kk >= I.ntz) {
            if (I.bc_z != 0) continue;
            kk = (kk + I.ntz) % I.ntz;
        }
        for (int dj = -1; dj <= 1; dj++) {
            int jj = tj + dj;
            if (jj < 0 || jj >= I.nty) {
                if (I.bc_y != 0) continue;
                jj = (jj + I.nty) % I.nty;
            }
            for (int di = -1; di <= 1; di++) {
                const int nzc = (di != 0) + (dj != 0) + (dk != 0);
                if (kind == 0 && nzc > 1) continue;
                if (kind == 1 && dk != 0 && (di != 0 || dj != 0)) continue;
                int ii = ti + di;
                if (ii < 0 || ii >= I.ntx) {
                    if (I.bc_x != 0) continue;
                    ii = (ii + I.ntx) % I.ntx;
                }
                r |= N[ii + jj * I.ntx + kk * I.ntx * I.nty];
            }
        }
    }
    return r != 0;
}

// 2D tile grid: neighbour columns/rows resolved once (-1 = clamped away).
struct Nb2 {
    int xm, xp, ym, yp;  // tile-grid column / row indices
};

__device__ __forceinline__ Nb2 nb2_of(int ti, int tj, int ntx, int nty, int bcx, int bcy) {
    Nb2 n;
    n.xm = ti > 0 ? ti - 1 : (bcx == 0 ? ntx - 1 : -1);
    n.xp = ti < ntx - 1 ? ti + 1 : (bcx == 0 ? 0 : -1);
    n.ym = tj > 0 ? tj - 1 : (bcy == 0 ? nty - 1 : -1);
    n.yp = tj < nty - 1 ? tj + 1 : (bcy == 0 ? 0 : -1);
    return n;
}

__device__ __forceinline__ int dilated2d(const uint8_t* N, int ti, int tj, int ntx, const Nb2& n, int kind) {
    const int row = tj * ntx;
    int r = N[ti + row];
    if (n.xm >= 0) r |= N[n.xm + row];
    if (n.xp >= 0) r |= N[n.xp + row];
    if (n.ym >= 0) r |= N[ti + n.ym * ntx];
    if (n.yp >= 0) r |= N[ti + n.yp * ntx];
    if (kind == 1) {
        if (n.ym >= 0 && n.xm >= 0) r |= N[n.xm + n.ym * ntx];
        if (n.ym >= 0 && n.xp >= 0) r |= N[n.xp + n.ym * ntx];
        if (n.yp >= 0 && n.xm >= 0) r |= N[n.xm + n.yp * ntx];
        if (n.yp >= 0 && n.xp >= 0) r |= N[n.xp + n.yp * ntx];
    }
    return r != 0;
}

template <int NT>
__device__ __forceinline__ int block_exclusive_scan(int v, int* lds, int* total) {
    constexpr int NW = NT / 64;
    int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    int incl = v;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        int y = __shfl_up(incl, off, 64);
        if (lane >= off) incl += y;
    }
    if (lane == 63) lds[wid] = incl;
    __syncthreads();
    int base = 0, tot = 0;
#pragma unroll
    for (int w = 0; w < NW; w++) {
        int c = lds[w];
        if (w < wid) base += c;
        tot += c;
    }
    __syncthreads();
    *total = tot;
    return base + incl - v;
}

constexpr int kFinNT = 256;
constexpr int kSegWords = kSegTiles / 64;

// Per-instance end-of-step update from the summed segment partials.
__device__ __forceinline__ void finish_instance(const Inst& J, int64_t listlen, unsigned long long c0,
                                                unsigned long long c1, unsigned long long cells, const double* cs,
                                                double vmax, int64_t step, int64_t updates, const double* scal_in,
                                                int rotate) {
    J.ctr[2] = listlen;
    if (!rotate) return;
    if (step < J.hist_cap) {
        J.hist[2 * step] = (int32_t)c0;
        J.hist[2 * step + 1] = J.nsets > 1 ? (int32_t)c1 : 0;
    }
    J.ctr[0] = step + 1;
    J.ctr[1] = updates + (int64_t)cells;
    for (int f = 0; f < J.ncons; f++) J.scal[8 + f] = scal_in[8 + f] + cs[f] / J.ntotal;
    const double used = scal_in[1];
    J.scal[2] = scal_in[2] + used;
    J.scal[3] = used;
    if (J.model == 5 && J.dt_mode == 1) J.scal[1] = J.cfl * (J.dx / vmax);
}

// LDS of the finalize body.
struct FinLds {
    uint64_t s_cur[2][kSegWords];
    uint64_t s_par[2][kSegWords];   // buffer parity words the next step sees
    uint64_t s_new[2][kSegWords];
    int s_pre[kSegWords + 1];
    double sd[5];
    int si[6];
    unsigned long long s_cnt[3];
    double s_cs[kMaxFields];
    double s_scal[16];
    int64_t s_step;
    int64_t s_upd;
    int s_last;
};

template <int NT>
__device__ __forceinline__ void finalize_body(const Inst* __restrict__ insts, const Seg* __restrict__ segs,
                                              int segi, int nseg, const InstSegs* __restrict__ isegs, int n_inst,
                                              SegPart* __restrict__ parts, int32_t* __restrict__ counts,
                                              int32_t* __restrict__ offsets, uint32_t* __restrict__ done,
                                              int rotate, FinLds& F) {
    static_assert(NT >= 64 && NT <= 256 && kSegTiles % NT == 0, "finalize layout");
    constexpr int KT = kSegTiles / NT;  // tiles per thread
    constexpr int WPR = NT / 64;        // words per round
    auto& s_cur = F.s_cur;
    auto& s_new = F.s_new;
    auto& s_par = F.s_par;
    auto& s_pre = F.s_pre;
    auto& sd = F.sd;
    auto& si = F.si;
    auto& s_cnt = F.s_cnt;
    auto& s_cs = F.s_cs;
    auto& s_scal = F.s_scal;
    auto& s_step = F.s_step;
    auto& s_upd = F.s_upd;
    auto& s_last = F.s_last;

    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    const Seg sg = segs[segi];
    const Inst& I = insts[sg.inst];
    const int t0 = sg.tile0, nts = sg.ntile;
    const int w0 = t0 >> 6, nws = (nts + 63) >> 6;
    const bool regular = I.mode == 0;
    const int nsets = I.nsets;
    const int nt = I.nt;
    const int kind = I.model == 5 ? 1 : 0;
    const bool solo = nseg == 1;

    // ---- prologue: every independent global load of the kernel, issued together
    uint64_t cur = 0ull, parw = 0ull;
    const int ws = tid / kSegWords, wj = tid % kSegWords;
    const bool wthread = tid < 2 * kSegWords && wj < nws && ws < nsets;
    static_assert(2 * kSegWords <= 64, "word threads fit in one wave");
    if (wthread) {
        cur = regular ? full_word(w0 + wj, nt) : I.a_cur[ws][w0 + wj];
        parw = I.par[ws][w0 + wj];
    }
    int a0[KT], a1[KT];
    {
        const bool flat = I.ntz == 1;
        const int ntx = I.ntx, nty = I.nty;
#pragma unroll
        for (int i = 0; i < KT; i++) {
            const int k = i * NT + tid;
            const int t = t0 + k;
            a0[i] = 0;
            a1[i] = 0;
            if (k >= nts || !rotate || regular) continue;
            if (flat) {
                const int tj = t / ntx, ti = t - tj * ntx;
                const Nb2 n = nb2_of(ti, tj, ntx, nty, I.bc_x, I.bc_y);
                a0[i] = dilated2d(I.nflag[0], ti, tj, ntx, n, kind);
                if (nsets > 1) a1[i] = dilated2d(I.nflag[1], ti, tj, ntx, n, kind);
            } else {
                a0[i] = dilated3d(I, I.nflag[0], t, kind);
                if (nsets > 1) a1[i] = dilated3d(I, I.nflag[1], t, kind);
            }
        }
    }
    if (solo && tid < 16) s_scal[tid] = I.scal[tid];
    if (solo && tid == 16) s_step = I.ctr[0];
    if (solo && tid == 17) s_upd = I.ctr[1];
    if (tid < 3) s_cnt[tid] = 0ull;
    if (wthread) {
        s_cur[ws][wj] = cur;
        s_par[ws][wj] = rotate ? (parw ^ cur) : parw;
    } else if (tid < 2 * kSegWords) {
        s_cur[ws][wj] = 0ull;
        s_par[ws][wj] = 0ull;
    }
    __syncthreads();

    if (rotate) {
        // (a) statistics of the executed step, parity flip
        if (wthread) I.par[ws][w0 + wj] = parw ^ cur;
        if (tid < nws) {
            const uint64_t u = s_cur[0][tid] | s_cur[1][tid];
            unsigned long long cells = 0;
            if (I.full_tiles) {
                cells = (unsigned long long)__popcll(u) * (unsigned long long)(I.tx * I.ty * I.tz);
            } else {
                uint64_t m = u;
                while (m) {
                    const int b = __ffsll((unsigned long long)m) - 1;
                    cells += (unsigned long long)tile_cells(I, (w0 + tid) * 64 + b);
                    m &= m - 1ull;
                }
            }
            atomicAdd(&s_cnt[0], (unsigned long long)__popcll(s_cur[0][tid]));
            atomicAdd(&s_cnt[1], (unsigned long long)__popcll(s_cur[1][tid]));
            atomicAdd(&s_cnt[2], cells);
        }
        // conserved change sums of the tiles that ran (sleeping tiles contribute 0)
        for (int f = 0; f < I.ncons; f++) {
            double csum = 0.0;
#pragma unroll
            for (int i = 0; i < KT; i++) {
                const int k = i * NT + tid;
                if (k < nts && ((s_cur[0][k >> 6] >> (k & 63)) & 1ull)) csum += I.tpart[(int64_t)f * nt + t0 + k];
            }
            const double tot = block_sum<NT>(csum, sd);
            if (tid == 0) s_cs[f] = tot;
        }
    }

    // (b) next awake set, one ballot per 64-tile word
#pragma unroll
    for (int i = 0; i < KT; i++) {
        const int k = i * NT + tid;
        int b0 = a0[i], b1 = a1[i];
        if (k < nts && (regular || !rotate)) {
            b0 = regular ? 1 : (int)((s_cur[0][k >> 6] >> (k & 63)) & 1ull);
            b1 = regular ? (nsets > 1) : (int)((s_cur[1][k >> 6] >> (k & 63)) & 1ull);
        }
        const uint64_t w0b = __ballot(b0);
        const uint64_t w1b = __ballot(b1);
        if (lane == 0) {
            s_new[0][i * WPR + wid] = w0b;
            s_new[1][i * WPR + wid] = w1b;
        }
    }
    __syncthreads();
    if (rotate && !regular && wthread) I.a_cur[ws][w0 + wj] = s_new[ws][wj];
    // KH: max wave speed the next step's CFL sees (fresh for awake tiles, stale otherwise)
    double vmax = 0.0;
    if (rotate && I.model == 5 && I.dt_mode == 1) {
        for (int k = tid; k < nts; k += NT) {
            const int a = (int)((s_new[0][k >> 6] >> (k & 63)) & 1ull);
            vmax = fmax(vmax, a ? I.tspeed[t0 + k] : I.tspeed[nt + t0 + k]);
        }
        vmax = block_max<NT>(vmax, sd);
    }

    // (c) compaction of the union set of the segment
    if (tid < kSegWords) {
        // exclusive prefix of the word popcounts (16 lanes of wave 0)
        const int c = __popcll(s_new[0][tid] | s_new[1][tid]);
        int incl = c;
#pragma unroll
        for (int off = 1; off < kSegWords; off <<= 1) {
            const int y = __shfl_up(incl, off, 64);
            if (tid >= off) incl += y;
        }
        s_pre[tid] = incl - c;
        if (tid == kSegWords - 1) s_pre[kSegWords] = incl;
    }
    __syncthreads();
    const bool flat2 = I.ntz == 1;
#pragma unroll
    for (int i = 0; i < KT; i++) {
        const int k = i * NT + tid;
        if (k >= nts) break;
        const int j = k >> 6, b = k & 63;
        const uint64_t u = s_new[0][j] | s_new[1][j];
        if (!((u >> b) & 1ull)) continue;
        const int pos = t0 + s_pre[j] + __popcll(u & ((1ull << b) - 1ull));
        I.seg[pos] = t0 + k;
        if (flat2) {
            // neighbour record: resolved here when all nine owners are in this segment
            uint64_t rec = kRecValid;
            for (int d = 0; d < 9; d++) {
                const int o = nb_owner2d(I, t0 + k, I.tx, d) - t0;
                if (o < 0 || o >= nts) {
                    rec = 0ull;
                    break;
                }
                for (int s2 = 0; s2 < nsets; s2++) {
                    rec |= ((s_par[s2][o >> 6] >> (o & 63)) & 1ull) << (s2 * 18 + d);
                    rec |= ((s_new[s2][o >> 6] >> (o & 63)) & 1ull) << (s2 * 18 + 9 + d);
                }
            }
            I.nbrec[pos] = rec;
        }
    }

    // (d) per-instance reductions and merged offsets
    if (solo) {
        if (tid == 0) {
            double cs[kMaxFields] = {s_cs[0], s_cs[1], s_cs[2], s_cs[3]};
            finish_instance(I, s_pre[kSegWords], s_cnt[0], s_cnt[1], s_cnt[2], cs, vmax, s_step, s_upd, s_scal,
                            rotate);
            offsets[0] = 0;
            offsets[1] = s_pre[kSegWords];
        }
        return;
    }
    if (tid == 0) {
        counts[segi] = s_pre[kSegWords];
        if (rotate) {
            parts[segi].c0 = s_cnt[0];
            parts[segi].c1 = s_cnt[1];
            parts[segi].cells = s_cnt[2];
            for (int f = 0; f < kMaxFields; f++) parts[segi].csum[f] = f < I.ncons ? s_cs[f] : 0.0;
            parts[segi].vmax = vmax;
        }
        __threadfence();
        const unsigned prev = atomicAdd(done, 1u);
        s_last = (prev == (unsigned)(nseg - 1));
    }
    __syncthreads();
    if (!s_last) return;
    __threadfence();
    for (int i = tid; i < n_inst; i += NT) {
        const Inst& J = insts[i];
        const InstSegs r = isegs[i];
        int64_t listlen = 0;
        unsigned long long c0 = 0, c1 = 0, cells = 0;
        double cs[kMaxFields] = {0.0, 0.0, 0.0, 0.0};
        double vm = 0.0;
        for (int q = r.first; q < r.first + r.count; q++) {
            listlen += __hip_atomic_load(&counts[q], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (rotate) {
                c0 += parts[q].c0;
                c1 += parts[q].c1;
                cells += parts[q].cells;
                for (int f = 0; f < J.ncons; f++) cs[f] += parts[q].csum[f];
                vm = fmax(vm, parts[q].vmax);
            }
        }
        double sc[16];
        for (int k = 0; k < 16; k++) sc[k] = J.scal[k];
        finish_instance(J, listlen, c0, c1, cells, cs, vm, J.ctr[0], J.ctr[1], sc, rotate);
    }
    int running = 0;
    for (int base = 0; base < nseg; base += NT) {
        const int i = base + tid;
        const int v = i < nseg ? __hip_atomic_load(&counts[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : 0;
        int tot;
        const int ex = block_exclusive_scan<NT>(v, si, &tot);
        if (i < nseg) offsets[i] = running + ex;
        running += tot;
    }
    if (tid == 0) {
        offsets[nseg] = running;
        *done = 0u;
    }
}


__global__ __launch_bounds__(kFinNT) void finalize_kernel(const Inst* __restrict__ insts,
                                                          const Seg* __restrict__ segs, int nseg,
                                                          const InstSegs* __restrict__ isegs, int n_inst,
                                                          SegPart* __restrict__ parts,
                                                          int32_t* __restrict__ counts,
                                                          int32_t* __restrict__ offsets,
                                                          uint32_t* __restrict__ done, int rotate) {
    __shared__ FinLds F;
    finalize_body<kFinNT>(insts, segs, blockIdx.x, nseg, isegs, n_inst, parts, counts, offsets, done, rotate, F);
}

// Current value of one field: the buffer the tile parity selects, minus the
// conserved shift accumulated since the tile last ran.
__global__ void materialize_kernel(const Inst* __restrict__ insts, int slot, int field, double* __restrict__ dst) {
    const Inst& I = insts[slot];
    const int64_t n = (int64_t)I.nx * I.ny * I.nz;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        dst[i] = current_value(I, field, i);
}

// After an upload: both buffers hold the new values, so the shift epochs of
// every tile restart at the current cumulative shift.
__global__ void reset_epochs_kernel(const Inst* __restrict__ insts, int slot) {
    const Inst& I = insts[slot];
    for (int f = 0; f < I.ncons; f++) {
        const double S = I.scal[8 + f];
        for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < 2 * I.nt; i += gridDim.x * blockDim.x)
            I.s_tile[(int64_t)f * 2 * I.nt + i] = S;
    }
}

// Non-finite check over the current values of all fields.
__global__ void finite_kernel(const Inst* __restrict__ insts, int slot, int* __restrict__ bad) {
    const Inst& I = insts[slot];
    const int64_t n = (int64_t)I.nx * I.ny * I.nz;
    for (int f = 0; f < I.nfields; f++) {
        const int set = I.field_set[f];
        for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
            const int x = (int)(i % I.nx);
            const int y = (int)((i / I.nx) % I.ny);
            const int z = (int)(i / ((int64_t)I.nx * I.ny));
            const int t = x / I.tx + (y / I.ty) * I.ntx + (z / I.tz) * I.ntx * I.nty;
            const double v = I.buf[bit_of(I.par[set], t)][f][i];
            if (!isfinite(v)) atomicOr(bad, 1);
        }
    }
}

}  // namespace dba
