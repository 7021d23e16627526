// k_finalize.h - end-of-step bookkeeping and active-tile compaction.
//
// Replaces the per-block tail of the reference step: Wakeup_Next clears the
// flags of the blocks it ran (example_Dendritic_growth/kernel.cu:526-531) and
// Wakeup_Neighbor sets the next-step flags (kernel.cu:577-634). Here the
// dilation already happened inside the step kernel (atomicOr into a_next), so
// this kernel only
//   (a) records the step (history, counters), flips the buffer parity of the
//       tiles that ran, and for conserved DBA turns the per-tile change sums
//       into the global mean shift (example_Reactive_ModelB/Conserved/kernel.cu:1152-1164);
//   (b) for KH with a CFL time step, reduces the per-tile maximum wave speed
//       into the next dt;
//   (c) rotates a_next into a_cur and compacts it into an ascending worklist
//       (bitmap words are wave ballots; popcount + workgroup prefix sum, so
//       the order equals copy_if over tile indices);
//   (d) the last workgroup to finish turns the per-instance counts into the
//       merged worklist offsets.
// One workgroup per instance.
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

template <int NT>
__global__ __launch_bounds__(NT) void finalize_kernel(const Inst* __restrict__ insts, int n_inst,
                                                      int32_t* __restrict__ counts,
                                                      int32_t* __restrict__ offsets,
                                                      uint32_t* __restrict__ done, int rotate) {
    __shared__ double sd[NT / 64 + 1];
    __shared__ int si[NT / 64 + 2];
    __shared__ int s_last;
    __shared__ unsigned long long s_cnt[3];
    const int slot = blockIdx.x;
    const Inst& I = insts[slot];
    const int tid = threadIdx.x;
    const int nw = I.nwords;
    const bool regular = I.mode == 0;
    const int nsets = I.nsets;

    if (tid < 3) s_cnt[tid] = 0ull;
    __syncthreads();

    if (rotate) {
        // (a) statistics of the executed step, parity flip
        unsigned long long c0 = 0, c1 = 0, cells = 0;
        for (int w = tid; w < nw; w += NT) {
            uint64_t cur0 = regular ? full_word(w, I.nt) : I.a_cur[0][w];
            uint64_t cur1 = nsets > 1 ? (regular ? full_word(w, I.nt) : I.a_cur[1][w]) : 0ull;
            c0 += __popcll(cur0);
            c1 += __popcll(cur1);
            I.par[0][w] ^= cur0;
            if (nsets > 1) I.par[1][w] ^= cur1;
            uint64_t u = cur0 | cur1;
            if (I.full_tiles) {
                cells += (unsigned long long)__popcll(u) * (unsigned long long)(I.tx * I.ty * I.tz);
            } else {
                uint64_t m = u;
                while (m) {
                    int b = __ffsll((unsigned long long)m) - 1;
                    cells += (unsigned long long)tile_cells(I, w * 64 + b);
                    m &= m - 1ull;
                }
            }
        }
        atomicAdd(&s_cnt[0], c0);
        atomicAdd(&s_cnt[1], c1);
        atomicAdd(&s_cnt[2], cells);
        // conserved: mean shift of each conserved field from the change sums of
        // the tiles that ran (sleeping tiles contribute 0)
        for (int f = 0; f < I.ncons; f++) {
            double csum = 0.0;
            const double* tp = I.tpart + (int64_t)f * I.nt;
            for (int w = tid; w < nw; w += NT) {
                uint64_t m = regular ? full_word(w, I.nt) : I.a_cur[0][w];
                while (m) {
                    int b = __ffsll((unsigned long long)m) - 1;
                    csum += tp[w * 64 + b];
                    m &= m - 1ull;
                }
            }
            double tot = block_sum<NT>(csum, sd);
            if (tid == 0) I.scal[8 + f] = I.scal[8 + f] + tot / I.ntotal;
        }
        __syncthreads();
        if (tid == 0) {
            int64_t step = I.ctr[0];
            if (step < I.hist_cap) {
                I.hist[2 * step] = (int32_t)s_cnt[0];
                I.hist[2 * step + 1] = nsets > 1 ? (int32_t)s_cnt[1] : 0;
            }
            I.ctr[0] = step + 1;
            I.ctr[1] += (int64_t)s_cnt[2];
            double used = I.scal[1];
            I.scal[2] = I.scal[2] + used;
            I.scal[3] = used;
        }
        __syncthreads();
        // (b) KH CFL time step from the per-tile speeds the next step will see
        if (I.model == 5 && I.dt_mode == 1) {
            double vmax = 0.0;
            for (int t = tid; t < I.nt; t += NT) {
                int nxt = regular ? 1 : bit_of(I.a_next[0], t);
                double v = nxt ? I.tspeed[t] : I.tspeed[I.nt + t];
                vmax = fmax(vmax, v);
            }
            vmax = block_max<NT>(vmax, sd);
            if (tid == 0) I.scal[1] = I.cfl * (I.dx / vmax);
            __syncthreads();
        }
    }

    // (c) compaction of the next step's set (union over sets) + rotation
    const int chunk = (nw + NT - 1) / NT;
    const int w0 = min(nw, tid * chunk), w1 = min(nw, w0 + chunk);
    int cnt = 0;
    for (int w = w0; w < w1; w++) {
        uint64_t u;
        if (regular) u = full_word(w, I.nt);
        else if (rotate) u = I.a_next[0][w] | (nsets > 1 ? I.a_next[1][w] : 0ull);
        else u = I.a_cur[0][w] | (nsets > 1 ? I.a_cur[1][w] : 0ull);
        cnt += __popcll(u);
    }
    int total;
    int pos = block_exclusive_scan<NT>(cnt, si, &total);
    for (int w = w0; w < w1; w++) {
        uint64_t u;
        if (regular) {
            u = full_word(w, I.nt);
        } else if (rotate) {
            uint64_t n0 = I.a_next[0][w];
            uint64_t n1 = nsets > 1 ? I.a_next[1][w] : 0ull;
            u = n0 | n1;
            I.a_cur[0][w] = n0;
            I.a_next[0][w] = 0ull;
            if (nsets > 1) {
                I.a_cur[1][w] = n1;
                I.a_next[1][w] = 0ull;
            }
        } else {
            u = I.a_cur[0][w] | (nsets > 1 ? I.a_cur[1][w] : 0ull);
        }
        while (u) {
            int b = __ffsll((unsigned long long)u) - 1;
            I.seg[pos++] = w * 64 + b;
            u &= u - 1ull;
        }
    }

    // (d) merged offsets, written by the last workgroup to finish
    if (tid == 0) {
        counts[slot] = total;
        I.ctr[2] = total;
        __threadfence();
        unsigned prev = atomicAdd(done, 1u);
        s_last = (prev == (unsigned)(n_inst - 1));
    }
    __syncthreads();
    if (!s_last) return;
    __threadfence();
    int running = 0;
    for (int base = 0; base < n_inst; base += NT) {
        int i = base + tid;
        int v = i < n_inst ? __hip_atomic_load(&counts[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : 0;
        int tot;
        int ex = block_exclusive_scan<NT>(v, si, &tot);
        if (i < n_inst) offsets[i] = running + ex;
        running += tot;
    }
    if (tid == 0) {
        offsets[n_inst] = running;
        *done = 0u;
    }
}

// Current value of one field: the buffer the tile parity selects, minus the
// conserved shift accumulated since the tile last ran.
__global__ void materialize_kernel(const Inst* __restrict__ insts, int slot, int field, double* __restrict__ dst) {
    const Inst& I = insts[slot];
    const int64_t n = (int64_t)I.nx * I.ny * I.nz;
    const int set = I.field_set[field];
    const bool cons = field < I.ncons;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        int x = (int)(i % I.nx);
        int y = (int)((i / I.nx) % I.ny);
        int z = (int)(i / ((int64_t)I.nx * I.ny));
        int t = x / I.tx + (y / I.ty) * I.ntx + (z / I.tz) * I.ntx * I.nty;
        int p = bit_of(I.par[set], t);
        double v = I.buf[p][field][i];
        if (cons) v -= (I.scal[8 + field] - I.s_tile[((int64_t)field * 2 + p) * I.nt + t]);
        dst[i] = v;
    }
}

// After an upload: both buffers hold the new values, so the shift epochs of
// every tile restart at the current cumulative shift.
__global__ void reset_epochs_kernel(const Inst* __restrict__ insts, int slot) {
    const Inst& I = insts[slot];
    for (int f = 0; f < I.ncons; f++) {
        double S = I.scal[8 + f];
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
            int x = (int)(i % I.nx);
            int y = (int)((i / I.nx) % I.ny);
            int z = (int)(i / ((int64_t)I.nx * I.ny));
            int t = x / I.tx + (y / I.ty) * I.ntx + (z / I.tz) * I.ntx * I.nty;
            double v = I.buf[bit_of(I.par[set], t)][f][i];
            if (!isfinite(v)) atomicOr(bad, 1);
        }
    }
}

}  // namespace dba
