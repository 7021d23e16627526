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

__device__ __forceinline__ int64_t tile_cells(CInst& I, int t) {
    int ti = t % I.ntx, tj = (t / I.ntx) % I.nty, tk = t / (I.ntx * I.nty);
    int wx = min(I.tx, I.nx - ti * I.tx), wy = min(I.ty, I.ny - tj * I.ty), wz = min(I.tz, I.nz - tk * I.tz);
    return (int64_t)wx * wy * wz;
}

// Awake flag of tile t for the next step: OR of the detector bytes over the
// tile and its neighbours (kind 0: faces; kind 1: 3x3 in-plane + z faces),
// wrapping where the instance is periodic and clamping elsewhere. A neighbour
// clamped away reads the tile itself (or, for an in-plane corner, the edge
// neighbour that exists), which the OR already holds: no load waits on a branch.
__device__ __forceinline__ int dilated3d(CInst& I, const gu8* N, int t, int kind) {
    const int ntx = I.ntx, nty = I.nty, ntz = I.ntz, nxy = ntx * nty;
    const int ti = t % ntx, tj = (t / ntx) % nty, tk = t / nxy;
    auto nb = [](int i, int d, int n, int bc) {
        const int j = i + d;
        return (j >= 0 && j < n) ? j : (bc == 0 ? (j + n) % n : i);
    };
    const int xm = nb(ti, -1, ntx, I.bc_x), xp = nb(ti, 1, ntx, I.bc_x);
    const int ym = nb(tj, -1, nty, I.bc_y) * ntx, yp = nb(tj, 1, nty, I.bc_y) * ntx;
    const int zm = nb(tk, -1, ntz, I.bc_z) * nxy, zp = nb(tk, 1, ntz, I.bc_z) * nxy;
    const int row = tj * ntx, pl = tk * nxy;
    int r = N[t] | N[xm + row + pl] | N[xp + row + pl] | N[ti + ym + pl] | N[ti + yp + pl] | N[ti + row + zm] |
            N[ti + row + zp];
    if (kind == 1) r |= N[xm + ym + pl] | N[xp + ym + pl] | N[xm + yp + pl] | N[xp + yp + pl];
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

__device__ __forceinline__ int dilated2d(const gu8* N, int ti, int tj, int ntx, const Nb2& n, int kind) {
    // A missing neighbour (zero-flux or wall boundary) reads the tile itself
    // instead, which the OR already holds (a missing corner falls back to the
    // edge neighbour that exists, or to the tile): no load depends on a branch,
    // so all of them are in flight at once instead of one after another.
    const int xm = n.xm >= 0 ? n.xm : ti, xp = n.xp >= 0 ? n.xp : ti;
    const int ym = n.ym >= 0 ? n.ym : tj, yp = n.yp >= 0 ? n.yp : tj;
    const int row = tj * ntx, rm = ym * ntx, rp = yp * ntx;
    int r = N[ti + row] | N[xm + row] | N[xp + row] | N[ti + rm] | N[ti + rp];
    if (kind == 1) r |= N[xm + rm] | N[xp + rm] | N[xm + rp] | N[xp + rp];
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

constexpr int kFinNT = 1024;  // one tile per thread for a full segment
constexpr int kSegWords = kSegTiles / 64;

// Per-instance end-of-step update from the summed segment partials.
__device__ __forceinline__ void finish_instance(CInst& J, int64_t listlen, unsigned long long c0,
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
    uint8_t s_pb[2][kSegTiles];     // per tile: parity the next step sees
    uint8_t s_ab[2][kSegTiles];     // per tile: awake in the next step
    int s_wcnt[kSegWords];          // per 64-tile word: tiles in the next worklist
    unsigned long long s_wran[kSegWords][3];  // per word: tiles that ran (set 0, set 1), their cells
    double sd[kFinNT / 64 + 1];
    double s_cs[kMaxFields];
    double s_scal[16];
    int64_t s_step;
    int64_t s_upd;
};

// One thread per tile of the segment, one wave per 64-tile bitmap word, so the
// whole bookkeeping needs a single workgroup barrier (plus the reductions of
// the conserved and CFL variants).
template <int NT>
__device__ __forceinline__ void finalize_body(const Inst* __restrict__ insts, const Seg* __restrict__ segs,
                                              int segi, int nseg, SegPart* __restrict__ parts,
                                              int32_t* __restrict__ counts, int32_t* __restrict__ offsets,
                                              int rotate, int one_inst, FinLds& F) {
    static_assert(NT == kSegTiles, "one thread per tile of a segment");
    auto& s_pb = F.s_pb;
    auto& s_ab = F.s_ab;
    auto& s_wcnt = F.s_wcnt;
    auto& s_wran = F.s_wran;
    auto& sd = F.sd;
    auto& s_cs = F.s_cs;
    auto& s_scal = F.s_scal;
    auto& s_step = F.s_step;
    auto& s_upd = F.s_upd;

    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    // one segment: instance 0, all of its tiles (no dependent load of the segment table)
    // one instance: segment segi holds its tiles from segi kSegTiles on, so the
    // segment table is not read (its load would come before the descriptor's)
    const bool one = nseg == 1 || one_inst;
    const int inst = one ? 0 : __builtin_amdgcn_readfirstlane(segs[segi].inst);
    CInst& I = ((CInst*)insts)[inst];
    const int t0 = nseg == 1 ? 0 : (one_inst ? segi * kSegTiles : __builtin_amdgcn_readfirstlane(segs[segi].tile0));
    const int nts = nseg == 1 ? I.nt
                              : (one_inst ? min(kSegTiles, I.nt - t0)
                                          : __builtin_amdgcn_readfirstlane(segs[segi].ntile));
    const int w0 = t0 >> 6, nws = (nts + 63) >> 6;
    const bool regular = I.mode == 0;
    const int nsets = I.nsets;
    const int nt = I.nt;
    const int kind = I.model == 5 ? 1 : 0;
    const bool solo = nseg == 1;
    const int k = tid, t = t0 + k;
    const bool valid = k < nts;
    const bool wvalid = wid < nws;  // this wave's word exists

    // ---- prologue: every independent global load, issued together
    uint64_t cur0 = 0ull, cur1 = 0ull, par0 = 0ull, par1 = 0ull;
    if (wvalid) {
        cur0 = regular ? full_word(w0 + wid, nt) : I.a_cur[0][w0 + wid];
        par0 = I.par[0][w0 + wid];
        if (nsets > 1) {
            cur1 = regular ? full_word(w0 + wid, nt) : I.a_cur[1][w0 + wid];
            par1 = I.par[1][w0 + wid];
        }
    }
    int a0 = 0, a1 = 0;
    if (valid && rotate && !regular) {
        if (I.ntz == 1) {
            const int ntx = I.ntx;
            const int tj = t / ntx, ti = t - tj * ntx;
            const Nb2 n = nb2_of(ti, tj, ntx, I.nty, I.bc_x, I.bc_y);
            a0 = dilated2d(I.nflag[0], ti, tj, ntx, n, kind);
            if (nsets > 1) a1 = dilated2d(I.nflag[1], ti, tj, ntx, n, kind);
        } else {
            a0 = dilated3d(I, I.nflag[0], t, kind);
            if (nsets > 1) a1 = dilated3d(I, I.nflag[1], t, kind);
        }
    }
    if (solo && tid < 16) s_scal[tid] = I.scal[tid];
    if (solo && tid == 16) s_step = I.ctr[0];
    if (solo && tid == 17) s_upd = I.ctr[1];

    // ---- this tile: ran in the executed step?  awake in the next?  parity next?
    const int ran0 = (int)((cur0 >> lane) & 1ull), ran1 = (int)((cur1 >> lane) & 1ull);
    int b0, b1;
    if (rotate && !regular) {
        b0 = a0;
        b1 = a1;
    } else {
        // regular: every tile every step; rebuild (rotate = 0): the current sets
        b0 = valid && ran0;
        b1 = valid && ran1;
    }
    const uint64_t pn0 = rotate ? (par0 ^ cur0) : par0;
    const uint64_t pn1 = rotate ? (par1 ^ cur1) : par1;
    const uint64_t nw0 = __ballot(b0);
    const uint64_t nw1 = __ballot(b1);
    const uint64_t uni = nw0 | nw1;
    s_ab[0][k] = (uint8_t)b0;
    s_ab[1][k] = (uint8_t)b1;
    s_pb[0][k] = (uint8_t)((pn0 >> lane) & 1ull);
    s_pb[1][k] = (uint8_t)((pn1 >> lane) & 1ull);
    if (rotate && wvalid && lane == 0) {
        I.par[0][w0 + wid] = pn0;
        if (nsets > 1) I.par[1][w0 + wid] = pn1;
        if (!regular) {
            I.a_cur[0][w0 + wid] = nw0;
            if (nsets > 1) I.a_cur[1][w0 + wid] = nw1;
        }
    }
    // per-word statistics of the executed step
    // field-set updates: set 0's fields over the tiles that ran it; the second
    // set's field (dendrite T) over the tiles that ran either set, since a
    // p-awake tile writes T too
    unsigned long long cells = 0ull;
    if (rotate && valid && (ran0 | ran1)) {
        const unsigned long long tc = I.full_tiles ? (unsigned long long)(I.tx * I.ty * I.tz)
                                                   : (unsigned long long)tile_cells(I, t);
        cells = tc * (unsigned long long)(nsets > 1 ? 1 + (ran0 ? 1 : 0) : 1);
    }
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) cells += __shfl_down(cells, off, 64);
    if (lane == 0) {
        s_wcnt[wid] = __popcll(uni);
        s_wran[wid][0] = (unsigned long long)__popcll(cur0);
        s_wran[wid][1] = (unsigned long long)__popcll(cur1);
        s_wran[wid][2] = cells;
    }
    // conserved change sums of the tiles that ran (sleeping tiles contribute 0)
    if (rotate) {
        for (int f = 0; f < I.ncons; f++) {
            const double csum = (valid && ran0) ? I.tpart[(int64_t)f * nt + t] : 0.0;
            const double tot = block_sum<NT>(csum, sd);
            if (tid == 0) s_cs[f] = tot;
        }
    }
    // KH: max wave speed the next step's CFL sees (fresh for awake tiles, stale otherwise)
    double vmax = 0.0;
    if (rotate && I.model == 5 && I.dt_mode == 1) {
        if (valid) vmax = b0 ? I.tspeed[t] : I.tspeed[nt + t];
        vmax = block_max<NT>(vmax, sd);
    }
    __syncthreads();

    // ---- compaction of the union set (ascending tiles) and the neighbour records
    int base = 0, listlen = 0;
    for (int w = 0; w < nws; w++) {
        const int c = s_wcnt[w];
        if (w < wid) base += c;
        listlen += c;
    }
    if ((uni >> lane) & 1ull) {
        const int pos = t0 + base + __popcll(uni & ((1ull << lane) - 1ull));
        I.seg[pos] = t;
        if (I.ntz == 1) {
            // resolved here when all nine owners are in this segment
            uint64_t rec = kRecValid;
            const int tj = t / I.ntx, ti = t - tj * I.ntx;
            for (int d = 0; d < 9; d++) {
                const int o = nb_owner2d(I, ti, tj, d) - t0;
                if (o < 0 || o >= nts) {
                    rec = 0ull;
                    break;
                }
                for (int s2 = 0; s2 < nsets; s2++) {
                    rec |= (uint64_t)s_pb[s2][o] << (s2 * 18 + d);
                    rec |= (uint64_t)s_ab[s2][o] << (s2 * 18 + 9 + d);
                }
            }
            I.nbrec[pos] = rec;
        }
    }
    unsigned long long c0 = 0, c1 = 0, ccells = 0;
    if (tid == 0) {
        for (int w = 0; w < nws; w++) {
            c0 += s_wran[w][0];
            c1 += s_wran[w][1];
            ccells += s_wran[w][2];
        }
    }

    // ---- per-instance reductions and merged offsets
    if (solo) {
        if (tid == 0) {
            double cs[kMaxFields] = {s_cs[0], s_cs[1], s_cs[2], s_cs[3]};
            finish_instance(I, listlen, c0, c1, ccells, cs, vmax, s_step, s_upd, s_scal, rotate);
            offsets[0] = 0;
            offsets[1] = listlen;
        }
        return;
    }
    if (tid == 0) {
        counts[segi] = listlen;
        if (rotate) {
            parts[segi].c0 = c0;
            parts[segi].c1 = c1;
            parts[segi].cells = ccells;
            for (int f = 0; f < kMaxFields; f++) parts[segi].csum[f] = f < I.ncons ? s_cs[f] : 0.0;
            parts[segi].vmax = vmax;
        }
    }
}

// Several segments: one workgroup merges the segment partials after the
// segment workgroups (the kernel boundary makes their writes visible; a
// last-workgroup-done counter would need an agent-scope fence in every
// segment workgroup, and each such fence writes back and invalidates its XCD's
// L2). Per instance, one wave: lanes stride over the instance's segments, then a
// fixed-shape shuffle tree (the same order on every run, and for a one-segment
// instance exactly that segment's partial). Then the merged list offsets.
template <int NT>
__device__ __forceinline__ void finalize_merge_body(const Inst* __restrict__ insts, int nseg,
                                                    const InstSegs* __restrict__ isegs, int n_inst,
                                                    const SegPart* __restrict__ parts,
                                                    const int32_t* __restrict__ counts,
                                                    int32_t* __restrict__ offsets, int rotate, int* si) {
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    // One workgroup: the instances (a wave each, NT/64 at a time), then the
    // offsets scan. Several: workgroup 0 scans, workgroup b >= 1 takes
    // instances [(b-1) NT/64, b NT/64), all at once.
    const bool one = gridDim.x == 1;
    int i_lo = 0, i_hi = n_inst;  // this workgroup's instances
    if (!one) {
        i_lo = blockIdx.x == 0 ? 0 : ((int)blockIdx.x - 1) * (NT / 64);
        i_hi = blockIdx.x == 0 ? 0 : min(n_inst, i_lo + NT / 64);
    }
    // the offsets scan's first counts, loaded with the reductions' inputs
    const int v_first = tid < nseg ? counts[tid] : 0;
    for (int i0 = i_lo; i0 < i_hi; i0 += NT / 64) {
        const int i = i0 + wid;
        if (i >= i_hi) break;
        CInst& J = ((CInst*)insts)[i];
        const InstSegs r = n_inst == 1 ? InstSegs{0, nseg} : isegs[i];
        // the instance's scalars and counters, loaded ahead of the reduction
        double sc[16];
        int64_t step = 0, upd = 0;
        if (lane == 0) {
            for (int kk = 0; kk < 16; kk++) sc[kk] = J.scal[kk];
            step = J.ctr[0];
            upd = J.ctr[1];
        }
        int64_t ll = 0;
        unsigned long long q0 = 0, q1 = 0, qc = 0;
        double cs[kMaxFields] = {0.0, 0.0, 0.0, 0.0};
        double vm = 0.0;
        const int ncs = J.ncons;
        for (int q = r.first + lane; q < r.first + r.count; q += 64) {
            ll += counts[q];
            if (rotate) {
                q0 += parts[q].c0;
                q1 += parts[q].c1;
                qc += parts[q].cells;
                for (int f = 0; f < ncs; f++) cs[f] += parts[q].csum[f];
                vm = fmax(vm, parts[q].vmax);
            }
        }
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) {
            ll += __shfl_down(ll, off, 64);
            q0 += __shfl_down(q0, off, 64);
            q1 += __shfl_down(q1, off, 64);
            qc += __shfl_down(qc, off, 64);
            for (int f = 0; f < kMaxFields; f++) cs[f] += __shfl_down(cs[f], off, 64);
            vm = fmax(vm, __shfl_down(vm, off, 64));
        }
        if (lane == 0) finish_instance(J, ll, q0, q1, qc, cs, vm, step, upd, sc, rotate);
    }
    if (!one && blockIdx.x != 0) return;
    int running = 0;
    for (int b = 0; b < nseg; b += NT) {
        const int i = b + tid;
        const int v = b == 0 ? v_first : (i < nseg ? counts[i] : 0);
        int tot;
        const int ex = block_exclusive_scan<NT>(v, si, &tot);
        if (i < nseg) offsets[i] = running + ex;
        running += tot;
    }
    if (tid == 0) offsets[nseg] = running;
}


__global__ __launch_bounds__(kFinNT) void finalize_kernel(const Inst* __restrict__ insts,
                                                          const Seg* __restrict__ segs, int nseg,
                                                          SegPart* __restrict__ parts,
                                                          int32_t* __restrict__ counts,
                                                          int32_t* __restrict__ offsets, int rotate,
                                                          int one_inst) {
    __shared__ FinLds F;
    finalize_body<kFinNT>(insts, segs, blockIdx.x, nseg, parts, counts, offsets, rotate, one_inst, F);
}

__global__ __launch_bounds__(kFinNT) void finalize_merge_kernel(const Inst* __restrict__ insts, int nseg,
                                                                const InstSegs* __restrict__ isegs, int n_inst,
                                                                const SegPart* __restrict__ parts,
                                                                const int32_t* __restrict__ counts,
                                                                int32_t* __restrict__ offsets, int rotate) {
    __shared__ int si[kFinNT / 64 + 2];
    finalize_merge_body<kFinNT>(insts, nseg, isegs, n_inst, parts, counts, offsets, rotate, si);
}

// Instance and tile of a single-launch step's workgroup. With `deal`, workgroups
// are dealt to XCDs in contiguous ranges (the worklist kernels' deal), so halo
// rows a neighbour tile wrote last step are in this XCD's L2 (CH, every tile
// active); without, tile b goes to XCD b % 8, which spreads a sparse set of
// awake tiles over all XCDs (K93, whose awake tiles cluster in a few rows). A group of several
// instances or segments maps the dealt index through dmap (instance << 24 |
// tile, the instances' tiles in group order); one instance in one segment
// uses the index itself.
__device__ __forceinline__ CInst& dense_item(const Inst* insts, const int32_t* dmap, int* t, bool deal = true) {
    const int b = deal ? xcd_deal(blockIdx.x, gridDim.x, 0, gridDim.x) : (int)blockIdx.x;
    if (!dmap) {
        *t = b;
        return ((CInst*)insts)[0];
    }
    const uint32_t d = (uint32_t)dmap[b];
    *t = (int)(d & 0xFFFFFFu);
    return ((CInst*)insts)[d >> 24];
}

// The detector and parity bytes a single-launch step reads (step parity prev):
// the four arrays' pointers in one batch of scalar loads, then a select. The
// empty asm keeps the compiler from turning the select of two loaded pointers
// back into a load from a selected address, which put a dependent scalar load
// in front of each array's bytes.
__device__ __forceinline__ void dense_bytes(CInst& I, int prev, const gu8** F, const gu8** P) {
    const gu8* F0 = I.dF[0];
    const gu8* F1 = I.dF[1];
    const gu8* P0 = I.dP[0];
    const gu8* P1 = I.dP[1];
    asm volatile("" : "+s"(F0), "+s"(F1), "+s"(P0), "+s"(P1));
    *F = prev ? F1 : F0;
    *P = prev ? P1 : P0;
}

// The descriptor fields a single-launch step reads before its first barrier,
// as one batch of scalar loads at the start of the workgroup: the empty asm
// needs every value at that point, so the loads issue together and cost one
// round trip, where the compiler had placed them one dependent wait after another.
__device__ __forceinline__ void dense_prefetch(CInst& I, int nfield) {
    const int nt = I.nt, ntx = I.ntx, nty = I.nty, nx = I.nx, ny = I.ny, bcx = I.bc_x, bcy = I.bc_y, mode = I.mode;
    const gu8 *f0 = I.dF[0], *f1 = I.dF[1], *p0 = I.dP[0], *p1 = I.dP[1];
    const gi32* runs = I.druns;
    const gdouble *b00 = I.buf[0][0], *b10 = I.buf[1][0];
    const gdouble *b01 = nfield > 1 ? I.buf[0][1] : b00, *b11 = nfield > 1 ? I.buf[1][1] : b10;
    asm volatile("" ::"s"(nt), "s"(ntx), "s"(nty), "s"(nx), "s"(ny), "s"(bcx), "s"(bcy), "s"(mode), "s"(f0), "s"(f1),
                 "s"(p0), "s"(p1), "s"(runs), "s"(b00), "s"(b10), "s"(b01), "s"(b11));
}

// Neighbour table of a single-launch step (threads 0..8; caller synchronises):
// per set, buffer parity and awake bit of each 3x3 owner tile, from the bitmaps
// at the first step of a run, afterwards from the previous step's per-tile bytes
// (awake = the 5-point dilation of its detector bits, the finalize's rule).
// dF / dP hold, per step parity, nsets x nt bytes (set-major).
__device__ __forceinline__ void dense_nb_fill(NbTable& nb, CInst& I, int t, int first, int prev, bool regular,
                                              int nsets, bool cons) {
    const int d = threadIdx.x;
    if (d >= 9) return;
    const int nt = I.nt, ntx = I.ntx, tj = t / ntx, ti = t - tj * ntx;
    const int o = nb_owner2d(I, ti, tj, d);
    const int oj = o / ntx, oi = o - oj * ntx;
    nb.tile[d] = o;
    // every load first (both sets, the conserved epochs), then the table
    const double e0 = cons ? I.s_tile[o] : 0.0, e1 = cons ? I.s_tile[nt + o] : 0.0;
    int par[2] = {0, 0}, act[2] = {1, 1};
    if (first) {
        for (int s = 0; s < nsets; s++) par[s] = bit_of(I.par[s], o);
        if (!regular)
            for (int s = 0; s < nsets; s++) act[s] = bit_of(I.a_cur[s], o);
    } else {
        const Nb2 n = nb2_of(oi, oj, ntx, I.nty, I.bc_x, I.bc_y);
        const gu8 *F, *P;
        dense_bytes(I, prev, &F, &P);
        for (int s = 0; s < nsets; s++) par[s] = P[s * nt + o];
        if (!regular)
            for (int s = 0; s < nsets; s++) act[s] = dilated2d(F + s * nt, oi, oj, ntx, n, 0);
    }
    for (int s = 0; s < nsets; s++) {
        nb.par[s][d] = par[s];
        nb.act[s][d] = act[s];
    }
    // conserved: the epoch of the owner's current buffer (the caller subtracts
    // it from the step's cumulative shift once that is known)
    for (int f = 0; f < kMaxFields; f++) nb.shift[f][d] = (cons && f == 0) ? (par[0] ? e1 : e0) : 0.0;
}

// End of a run of n >= 1 single-launch steps: the state the worklist path and
// the host read, from the per-tile bytes of the last step. Awake bitmap = the
// dilation of the last detector bits, parity bitmap, detector bytes (a tile
// that did not fire holds 0, the invariant the finalize's dilation relies on),
// steps, cell updates, the awake count, and the simulated time advanced step by
// step as the finalize advances it. One thread per tile of a segment (the
// segment starts on a 64-tile word); returns, in every thread, the segment's
// cell updates and awake-tile count.
__device__ __forceinline__ void dense_canon_seg(CInst& I, int t0, int ntile, int n, unsigned long long* scells,
                                                int* scnt, unsigned long long* cells_out, int* cnt_out) {
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    const int nt = I.nt, ntx = I.ntx;
    const int t = t0 + tid;
    const bool valid = tid < ntile;
    const bool regular = I.mode == 0;
    const int last = n & 1;
    const int nsets = I.nsets;
    const int nw = (ntile + 63) >> 6;
    // every load first, unconditionally (a thread past the segment reads tile t0,
    // a one-set instance reads set 0 twice; neither result is used), so they are
    // in flight together instead of waited for one branch at a time
    const int tc = valid ? t : t0;
    const int tj = tc / ntx, ti = tc - tj * ntx;
    const Nb2 nbr = nb2_of(ti, tj, ntx, I.nty, I.bc_x, I.bc_y);
    const gu8 *F, *P;
    dense_bytes(I, last, &F, &P);
    const int r0 = I.druns[tc];
    int fv[2], pv[2], av[2];
    for (int s = 0; s < 2; s++) {
        const int so = s < nsets ? s * nt : 0;
        fv[s] = F[so + tc];
        pv[s] = P[so + tc];
        av[s] = regular ? 1 : dilated2d(F + so, ti, tj, ntx, nbr, 0);
    }
    const int r = valid ? r0 : 0;
    uint64_t wa = 0ull;
    for (int s = 0; s < nsets; s++) {
        const int f = valid ? fv[s] : 0, p = valid ? pv[s] : 0, a = valid ? av[s] : 0;
        if (valid) I.nflag[s][t] = (uint8_t)f;
        const uint64_t ws = __ballot(a), wp = __ballot(p);
        wa |= ws;
        if (lane == 0 && wid < nw) {
            I.par[s][(t0 >> 6) + wid] = wp;
            if (!regular) I.a_cur[s][(t0 >> 6) + wid] = ws;
        }
    }
    if (valid) I.druns[t] = 0;
    unsigned long long cells =
        valid ? (unsigned long long)r * (unsigned long long)(I.full_tiles ? I.tx * I.ty * I.tz : tile_cells(I, t))
              : 0ull;
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) cells += __shfl_down(cells, off, 64);
    if (lane == 0) {
        scells[wid] = cells;
        scnt[wid] = __popcll(wa);
    }
    __syncthreads();
    unsigned long long c = 0;
    int cnt = 0;
    for (int w = 0; w < nw; w++) {
        c += scells[w];
        cnt += scnt[w];
    }
    *cells_out = c;
    *cnt_out = cnt;
}

// The instance counters after n single-launch steps (one thread).
__device__ __forceinline__ void dense_canon_finish(CInst& I, int n, unsigned long long cells, int64_t cnt) {
    I.ctr[0] = I.ctr[0] + n;
    I.ctr[1] = I.ctr[1] + (int64_t)cells;
    I.ctr[2] = cnt;
    const double dt = I.scal[1];
    double tm = I.scal[2];
    for (int i = 0; i < n; i++) tm += dt;
    I.scal[2] = tm;
    I.scal[3] = dt;
}

// One instance in one segment: one workgroup does it all. Conserved (CH): the
// cumulative shift after the last step, from the last step's change sums in the
// finalize's order (block_sum over the tiles, 0 for tiles that did not run).
__global__ __launch_bounds__(kFinNT) void dense_canon_kernel(const Inst* __restrict__ insts, int n) {
    __shared__ unsigned long long scells[kFinNT / 64];
    __shared__ int scnt[kFinNT / 64];
    __shared__ double sd[kFinNT / 64 + 1];
    CInst& I = ((CInst*)insts)[0];
    unsigned long long c;
    int cnt;
    dense_canon_seg(I, 0, I.nt, n, scells, scnt, &c, &cnt);
    if (I.mode == 2) {
        const int last = n & 1, t = threadIdx.x;
        const double v = t < I.nt ? I.tpart[(int64_t)(1 + last) * I.nt + t] : 0.0;
        const double tot = block_sum<kFinNT>(v, sd);
        if (t == 0) I.scal[8] = I.scal[12 + last] + tot / I.ntotal;
    }
    if (threadIdx.x == 0) dense_canon_finish(I, n, c, cnt);
}

// Several segments or instances: one workgroup per segment writes its partial
// sums (parts[seg].cells, counts[seg]) ...
__global__ __launch_bounds__(kFinNT) void dense_canon_seg_kernel(const Inst* __restrict__ insts,
                                                                 const Seg* __restrict__ segs,
                                                                 SegPart* __restrict__ parts,
                                                                 int32_t* __restrict__ counts, int n) {
    __shared__ unsigned long long scells[kFinNT / 64];
    __shared__ int scnt[kFinNT / 64];
    const Seg sg = segs[blockIdx.x];
    CInst& I = ((CInst*)insts)[sg.inst];
    unsigned long long c;
    int cnt;
    dense_canon_seg(I, sg.tile0, sg.ntile, n, scells, scnt, &c, &cnt);
    if (threadIdx.x == 0) {
        parts[blockIdx.x].cells = c;
        counts[blockIdx.x] = cnt;
    }
}

// ... and one wave per instance sums them in segment order.
__global__ __launch_bounds__(kFinNT) void dense_canon_merge_kernel(const Inst* __restrict__ insts,
                                                                   const InstSegs* __restrict__ isegs, int n_inst,
                                                                   const SegPart* __restrict__ parts,
                                                                   const int32_t* __restrict__ counts, int n) {
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    for (int i = wid; i < n_inst; i += kFinNT / 64) {
        const InstSegs r = isegs[i];
        unsigned long long c = 0;
        int64_t cnt = 0;
        for (int q = r.first + lane; q < r.first + r.count; q += 64) {
            c += parts[q].cells;
            cnt += counts[q];
        }
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) {
            c += __shfl_down(c, off, 64);
            cnt += __shfl_down(cnt, off, 64);
        }
        if (lane == 0) dense_canon_finish(((CInst*)insts)[i], n, c, cnt);
    }
}

// Current value of one field: the buffer the tile parity selects, minus the
// conserved shift accumulated since the tile last ran.
__global__ void materialize_kernel(const Inst* __restrict__ insts, int slot, int field, double* __restrict__ dst) {
    CInst& I = ((CInst*)insts)[slot];
    const int64_t n = (int64_t)I.nx * I.ny * I.nz;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        dst[i] = current_value(I, field, i);
}

// After an upload: both buffers hold the new values, so the shift epochs of
// every tile restart at the current cumulative shift.
__global__ void reset_epochs_kernel(const Inst* __restrict__ insts, int slot) {
    CInst& I = ((CInst*)insts)[slot];
    for (int f = 0; f < I.ncons; f++) {
        const double S = I.scal[8 + f];
        for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < 2 * I.nt; i += gridDim.x * blockDim.x)
            I.s_tile[(int64_t)f * 2 * I.nt + i] = S;
    }
}

// Non-finite check over the current values of all fields.
__global__ void finite_kernel(const Inst* __restrict__ insts, int slot, int* __restrict__ bad) {
    CInst& I = ((CInst*)insts)[slot];
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
