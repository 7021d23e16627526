// k_droplets.h - droplet (coarsening) analysis of a field on the device: the
// rule of example_Reactive_ModelB/coarsening_analysis.cpp:48-170, which runs on
// the host over a downloaded 256^3 VTK file.
//
//   avg      = (max(0.5, max v) + min(0.5, min v)) / 2          (:118-140)
//   droplet  = a periodic face-connected component (6 neighbours in 3D, 4 in
//              2D) of the cells with !(v < avg) that holds a cell v > avg
//              (the flood fill starts only from cells > avg, :48-111)
//   volume   = number of cells v > avg                            (:145-156)
//
// Components come from a lock-free union-find over the cells: every cell is
// united with its +x, +y (+z) periodic neighbour, a union links the larger root
// under the smaller by compare-and-swap, so each final root is the smallest
// index of its component and the forest only ever points to smaller indices.
// Reads of the forest during the union pass are agent-scope atomics: they bypass
// the non-coherent per-CU cache, so a retry after a failed swap always sees the
// newer parent. Every failed swap means another union succeeded, so each lane's
// loop ends.
#pragma once
#include "common.h"

namespace dba {

constexpr int kCcBlocks = 1024;
constexpr int kCcNT = 256;

// part[2 b] = max, part[2 b + 1] = min of the field over block b's cells
// (seeded with 0.5 like the reference).
__global__ __launch_bounds__(kCcNT) void cc_minmax_kernel(const Inst* __restrict__ insts, int slot, int field,
                                                          double* __restrict__ part) {
    __shared__ double smx[kCcNT], smn[kCcNT];
    CInst& I = ((CInst*)insts)[slot];
    const int64_t n = (int64_t)I.nx * I.ny * I.nz;
    double mx = 0.5, mn = 0.5;
    for (int64_t i = (int64_t)blockIdx.x * kCcNT + threadIdx.x; i < n; i += (int64_t)kCcBlocks * kCcNT) {
        const double v = current_value(I, field, i);
        if (v > mx) mx = v;
        if (v < mn) mn = v;
    }
    smx[threadIdx.x] = mx;
    smn[threadIdx.x] = mn;
    __syncthreads();
    for (int s = kCcNT / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            smx[threadIdx.x] = fmax(smx[threadIdx.x], smx[threadIdx.x + s]);
            smn[threadIdx.x] = fmin(smn[threadIdx.x], smn[threadIdx.x + s]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = smx[0];
        part[2 * blockIdx.x + 1] = smn[0];
    }
}

// out[0] = avg from the block extremes (min and max are exact in any order)
__global__ void cc_threshold_kernel(const double* __restrict__ part, double* __restrict__ out) {
    if (threadIdx.x != 0) return;
    double mx = 0.5, mn = 0.5;
    for (int b = 0; b < kCcBlocks; b++) {
        if (part[2 * b] > mx) mx = part[2 * b];
        if (part[2 * b + 1] < mn) mn = part[2 * b + 1];
    }
    out[0] = (mx + mn) / 2.;
}

// cls = 2 (v > avg: seed and volume), 1 (!(v < avg): member only), 0; parent = self
__global__ __launch_bounds__(kCcNT) void cc_init_kernel(const Inst* __restrict__ insts, int slot, int field,
                                                        const double* __restrict__ avgp, uint8_t* __restrict__ cls,
                                                        int32_t* __restrict__ parent, int32_t* __restrict__ size,
                                                        int32_t* __restrict__ key) {
    CInst& I = ((CInst*)insts)[slot];
    const double avg = avgp[0];
    const int64_t n = (int64_t)I.nx * I.ny * I.nz;
    for (int64_t i = (int64_t)blockIdx.x * kCcNT + threadIdx.x; i < n; i += (int64_t)kCcBlocks * kCcNT) {
        const double v = current_value(I, field, i);
        cls[i] = v > avg ? 2 : (!(v < avg) ? 1 : 0);
        parent[i] = (int32_t)i;
        size[i] = 0;
        key[i] = INT32_MAX;
    }
}

__device__ __forceinline__ int32_t cc_load(int32_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Root of i with path halving; every pointer written is an ancestor of its cell.
__device__ __forceinline__ int32_t cc_find(int32_t* parent, int32_t i) {
    for (;;) {
        const int32_t p = cc_load(&parent[i]);
        if (p == i) return i;
        const int32_t g = cc_load(&parent[p]);
        if (g != p) __hip_atomic_store(&parent[i], g, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        i = g;
    }
}

__device__ __forceinline__ void cc_unite(int32_t* parent, int32_t a, int32_t b) {
    for (;;) {
        a = cc_find(parent, a);
        b = cc_find(parent, b);
        if (a == b) return;
        if (a < b) {
            const int32_t t = a;
            a = b;
            b = t;
        }
        int32_t expect = a;
        if (__hip_atomic_compare_exchange_strong(&parent[a], &expect, b, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                                 __HIP_MEMORY_SCOPE_AGENT))
            return;
    }
}

__global__ __launch_bounds__(kCcNT) void cc_union_kernel(const uint8_t* __restrict__ cls, int32_t* parent, int nx,
                                                         int ny, int nz) {
    const int64_t nxy = (int64_t)nx * ny, n = nxy * nz;
    for (int64_t i = (int64_t)blockIdx.x * kCcNT + threadIdx.x; i < n; i += (int64_t)kCcBlocks * kCcNT) {
        if (!cls[i]) continue;
        const int x = (int)(i % nx), y = (int)((i / nx) % ny), z = (int)(i / nxy);
        const int64_t yz = i - x;
        const int64_t jx = yz + (x + 1 == nx ? 0 : x + 1);
        const int64_t jy = x + (y + 1 == ny ? 0 : y + 1) * (int64_t)nx + z * nxy;
        const int64_t jz = x + y * (int64_t)nx + (z + 1 == nz ? 0 : z + 1) * nxy;
        if (cls[jx]) cc_unite(parent, (int32_t)i, (int32_t)jx);
        if (cls[jy]) cc_unite(parent, (int32_t)i, (int32_t)jy);
        if (nz > 1 && cls[jz]) cc_unite(parent, (int32_t)i, (int32_t)jz);
    }
}

// Final roots; per root the member count and its first seed (smallest index of
// a cell > avg); vpart[b] = seeds (the volume) counted by block b.
__global__ __launch_bounds__(kCcNT) void cc_count_kernel(const uint8_t* __restrict__ cls, int32_t* parent,
                                                         int32_t* size, int32_t* key, int64_t n,
                                                         int64_t* __restrict__ vpart) {
    __shared__ int64_t sv[kCcNT];
    int64_t vol = 0;
    for (int64_t i = (int64_t)blockIdx.x * kCcNT + threadIdx.x; i < n; i += (int64_t)kCcBlocks * kCcNT) {
        const int c = cls[i];
        if (!c) continue;
        const int32_t r = cc_find(parent, (int32_t)i);
        atomicAdd(&size[r], 1);
        if (c == 2) {
            atomicMin(&key[r], (int32_t)i);
            vol++;
        }
    }
    sv[threadIdx.x] = vol;
    __syncthreads();
    for (int s = kCcNT / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) sv[threadIdx.x] += sv[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) vpart[blockIdx.x] = sv[0];
}

// Droplets: roots holding a seed, appended as (first seed, size) pairs in any
// order (the host sorts by seed); counters[0] = droplets, counters[1] = volume.
__global__ __launch_bounds__(kCcNT) void cc_collect_kernel(const int32_t* __restrict__ size,
                                                           const int32_t* __restrict__ key, int64_t n,
                                                           const int64_t* __restrict__ vpart,
                                                           int64_t* __restrict__ counters,
                                                           int64_t* __restrict__ out, int64_t cap) {
    for (int64_t i = (int64_t)blockIdx.x * kCcNT + threadIdx.x; i < n; i += (int64_t)kCcBlocks * kCcNT) {
        if (key[i] == INT32_MAX || size[i] == 0) continue;
        const int64_t k = (int64_t)atomicAdd((unsigned long long*)&counters[0], 1ull);
        if (k < cap) {
            out[2 * k] = key[i];
            out[2 * k + 1] = size[i];
        }
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        int64_t v = 0;
        for (int b = 0; b < kCcBlocks; b++) v += vpart[b];
        counters[1] = v;
    }
}

// ---- the same analysis with every buffer in one caller-given region ----
// (dba_droplets_noalloc). The droplets come out in the reference's order from
// the device itself: first[s] = size of the droplet whose first seed is cell s
// (0 elsewhere), then an ordered compaction of first[] (contiguous chunks per
// workgroup, chunk offsets from the per-chunk counts), so the host neither
// allocates nor sorts.

// p[0..n) = v; workgroup 0 also zeroes counters[0..2)
__global__ __launch_bounds__(kCcNT) void cc_fill_kernel(int32_t* __restrict__ p, int64_t n, int32_t v,
                                                        int64_t* __restrict__ counters) {
    for (int64_t i = (int64_t)blockIdx.x * kCcNT + threadIdx.x; i < n; i += (int64_t)kCcBlocks * kCcNT) p[i] = v;
    if (blockIdx.x == 0 && threadIdx.x < 2) counters[threadIdx.x] = 0;
}

// first[key[r]] = size[r] for every root r holding a seed (first[] zeroed);
// counters[0] = droplets, counters[1] = volume (the seeds counted per block).
__global__ __launch_bounds__(kCcNT) void cc_mark_kernel(const int32_t* __restrict__ size,
                                                        const int32_t* __restrict__ key, int64_t n,
                                                        const int64_t* __restrict__ vpart,
                                                        int32_t* __restrict__ first,
                                                        int64_t* __restrict__ counters) {
    __shared__ int64_t sc[kCcNT];
    int64_t cnt = 0;
    for (int64_t i = (int64_t)blockIdx.x * kCcNT + threadIdx.x; i < n; i += (int64_t)kCcBlocks * kCcNT) {
        const int32_t k = key[i], sz = size[i];
        if (k == INT32_MAX || sz == 0) continue;
        first[k] = sz;
        cnt++;
    }
    sc[threadIdx.x] = cnt;
    __syncthreads();
    for (int s = kCcNT / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) sc[threadIdx.x] += sc[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0 && sc[0] > 0) atomicAdd((unsigned long long*)&counters[0], (unsigned long long)sc[0]);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        int64_t v = 0;
        for (int b = 0; b < kCcBlocks; b++) v += vpart[b];
        counters[1] = v;
    }
}

// Chunk of workgroup b: cells [b * chunk, min(n, (b + 1) * chunk)).
__device__ __forceinline__ void cc_chunk(int64_t n, int64_t chunk, int64_t* lo, int64_t* hi) {
    *lo = (int64_t)blockIdx.x * chunk;
    *hi = *lo + chunk < n ? *lo + chunk : n;
}

// bcount[b] = droplet starts (first[] != 0) in chunk b
__global__ __launch_bounds__(kCcNT) void cc_order_count_kernel(const int32_t* __restrict__ first, int64_t n,
                                                               int64_t chunk, int32_t* __restrict__ bcount) {
    __shared__ int32_t sc[kCcNT];
    int64_t lo, hi;
    cc_chunk(n, chunk, &lo, &hi);
    int32_t c = 0;
    for (int64_t i = lo + threadIdx.x; i < hi; i += kCcNT) c += first[i] != 0;
    sc[threadIdx.x] = c;
    __syncthreads();
    for (int s = kCcNT / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) sc[threadIdx.x] += sc[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) bcount[blockIdx.x] = sc[0];
}

// out[k] = the k-th droplet start's size in cell order, for k < cap
__global__ __launch_bounds__(kCcNT) void cc_order_write_kernel(const int32_t* __restrict__ first, int64_t n,
                                                               int64_t chunk, const int32_t* __restrict__ bcount,
                                                               int32_t* __restrict__ out, int64_t cap) {
    constexpr int NW = kCcNT / 64;
    __shared__ int64_t sb[kCcNT];
    __shared__ int32_t sw[NW];
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    int64_t lo, hi;
    cc_chunk(n, chunk, &lo, &hi);
    // this chunk's offset: the starts of every earlier chunk
    int64_t b = 0;
    for (int j = tid; j < (int)blockIdx.x; j += kCcNT) b += bcount[j];
    sb[tid] = b;
    __syncthreads();
    for (int s = kCcNT / 2; s > 0; s >>= 1) {
        if (tid < s) sb[tid] += sb[tid + s];
        __syncthreads();
    }
    int64_t base = sb[0];
    for (int64_t w0 = lo; w0 < hi; w0 += kCcNT) {
        const int64_t i = w0 + tid;
        const int32_t v = i < hi ? first[i] : 0;
        const uint64_t m = __ballot(v != 0);
        if (lane == 0) sw[wid] = __popcll(m);
        __syncthreads();
        int64_t pre = 0, tot = 0;
        for (int w = 0; w < NW; w++) {
            pre += w < wid ? sw[w] : 0;
            tot += sw[w];
        }
        const int64_t k = base + pre + __popcll(m & ((1ull << lane) - 1ull));
        if (v != 0 && k < cap) out[k] = v;
        base += tot;
        __syncthreads();
    }
}

}  // namespace dba
