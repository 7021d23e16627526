// common.h - device-visible instance descriptor and helpers shared by all
// step kernels of the DBA library (gfx950 / wave64).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace dba {

constexpr int kMaxFields = 4;
constexpr int kMaxAux = 32;

// One simulation instance as the kernels see it. Lives in device memory
// (an array of these per merged group); everything the step needs is here.
struct Inst {
    double* buf[2][kMaxFields];  // double-buffered prognostic fields (tile parity picks one)
    double* aux[kMaxAux];        // model state kept for sleeping-neighbour semantics
    uint64_t* a_cur[2];          // awake bitmap of the step being executed, per set
    uint64_t* a_next[2];         // awake bitmap of the next step (built by dilation atomics)
    uint64_t* par[2];            // per-tile buffer parity per set: bit=1 -> buf[1] is current
    double* tpart;               // conserved: per-tile change sums, [field][nt]
    double* tspeed;              // KH: per-tile max wave speed, [0..nt) fresh (current U), [nt..2nt) stale
    double* s_tile;              // conserved: per-tile epoch of the cumulative shift, [field][parity][nt]
    double* scal;                // [1] dt of the next step, [2] time, [3] last dt, [8+f] cumulative shift S_f
    int64_t* ctr;                // [0] steps done, [1] active cell updates, [2] worklist length
    int32_t* hist;               // per-step active tile counts, [2*step + set]
    int32_t* seg;                // this instance's compacted worklist (ascending tiles)
    int32_t hist_cap;
    int32_t model, mode;
    int32_t nx, ny, nz;
    int32_t tx, ty, tz;
    int32_t ntx, nty, ntz, nt, nwords;
    int32_t bc_x, bc_y, bc_z;
    int32_t stencil, aniso_j, dt_mode, metric_kind;
    int32_t nsets;               // 2 for dendrites (p and T sets), else 1
    int32_t nfields;             // prognostic fields
    int32_t ncons;               // conserved DBA: fields carrying a mean shift
    int32_t field_set[kMaxFields];  // activation set (and parity bitmap) each field follows
    int32_t full_tiles;          // 1 if every tile is complete (no partial edge tiles)
    double dx, dy, dz, dt;
    double thresh_a, thresh_b;
    double ntotal;               // number of cells (conserved mean)
    // derived constants, computed on the host with the same expressions as the oracle
    double invdx2, inv6dx2, inv2dx, dt_tau, alpha_pi, cos_jt0, sin_jt0, k_epsp;
    // model constants
    double tau, eps_bar, delta, alpha, gamma, theta0, latent_k, t_eq;
    double mobility, kappa, react_gamma, cbar;
    double chi_xy, chi_xz, chi_yz, chi_xb, chi_yb, chi_zb;
    double v_x, v_y, v_z, r_x, r_y, r_z, n_exp, m_exp;
    double epsx_sq, epsy_sq, epsz_sq, epsphi_sq;
    double mob_x, mob_y, mob_z, mob_phi;
    double k_eq, z_crit, p_gel, k_0;
    double gas_gamma, cfl, vol;
};

// Boundary index map: periodic wrap, or mirror about the boundary face.
__device__ __forceinline__ int map_index(int i, int n, int bc) {
    if (bc == 0) {
        if (i < 0) i += n;
        else if (i >= n) i -= n;
        return i;
    }
    if (i < 0) return -1 - i;
    if (i >= n) return 2 * n - 1 - i;
    return i;
}

__device__ __forceinline__ int bit_of(const uint64_t* bm, int t) {
    return (int)((bm[t >> 6] >> (t & 63)) & 1ull);
}

// Set tile t and its neighbours in the next-step bitmap (Wakeup_Neighbor).
// kind 0: face neighbours (4 in 2D, 6 in 3D); kind 1: 3x3 in-plane + z faces.
__device__ __forceinline__ void dilate_into(uint64_t* bm, const Inst& I, int t, int kind) {
    int ti = t % I.ntx;
    int tj = (t / I.ntx) % I.nty;
    int tk = t / (I.ntx * I.nty);
    for (int dk = -1; dk <= 1; dk++) {
        for (int dj = -1; dj <= 1; dj++) {
            for (int di = -1; di <= 1; di++) {
                int nzc = (di != 0) + (dj != 0) + (dk != 0);
                if (kind == 0 && nzc > 1) continue;
                if (kind == 1 && dk != 0 && (di != 0 || dj != 0)) continue;
                int ii = ti + di, jj = tj + dj, kk = tk + dk;
                if (ii < 0 || ii >= I.ntx) { if (I.bc_x != 0) continue; ii = (ii + I.ntx) % I.ntx; }
                if (jj < 0 || jj >= I.nty) { if (I.bc_y != 0) continue; jj = (jj + I.nty) % I.nty; }
                if (kk < 0 || kk >= I.ntz) { if (I.bc_z != 0) continue; kk = (kk + I.ntz) % I.ntz; }
                int tt = ii + jj * I.ntx + kk * I.ntx * I.nty;
                atomicOr((unsigned long long*)&bm[tt >> 6], 1ull << (tt & 63));
            }
        }
    }
}

// Workgroup-wide OR of a per-thread predicate (wave ballot, then LDS across waves).
template <int NT>
__device__ __forceinline__ int block_any(int pred, int* lds_scratch) {
    constexpr int NW = NT / 64;
    unsigned long long b = __ballot(pred);
    int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) lds_scratch[wid] = b != 0ull;
    __syncthreads();
    int r = 0;
#pragma unroll
    for (int w = 0; w < NW; w++) r |= lds_scratch[w];
    __syncthreads();
    return r;
}

// Deterministic workgroup sum: per-wave butterfly-free shift reduction in a
// fixed order, then wave totals added in wave order by thread 0.
template <int NT>
__device__ __forceinline__ double block_sum(double v, double* lds_scratch) {
    constexpr int NW = NT / 64;
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_down(v, off, 64);
    int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) lds_scratch[wid] = v;
    __syncthreads();
    double r = 0.0;
    if (threadIdx.x == 0) {
        for (int w = 0; w < NW; w++) r += lds_scratch[w];
        lds_scratch[NW] = r;
    }
    __syncthreads();
    r = lds_scratch[NW];
    __syncthreads();
    return r;
}

template <int NT>
__device__ __forceinline__ double block_max(double v, double* lds_scratch) {
    constexpr int NW = NT / 64;
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v = fmax(v, __shfl_down(v, off, 64));
    int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) lds_scratch[wid] = v;
    __syncthreads();
    double r = lds_scratch[0];
#pragma unroll
    for (int w = 1; w < NW; w++) r = fmax(r, lds_scratch[w]);
    __syncthreads();
    return r;
}

// Map a grid-stride round-robin block index so that consecutive worklist
// items land on the same XCD (workgroups are dealt to the 8 XCDs in turn).
__device__ __forceinline__ int xcd_remap(int b, int G) {
    if ((G & 7) != 0) return b;
    return (b & 7) * (G >> 3) + (b >> 3);
}

// Worklist item -> instance slot (offsets is the exclusive prefix, offsets[n] = total).
__device__ __forceinline__ int find_slot(const int32_t* offsets, int n, int item) {
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (offsets[mid] <= item) lo = mid; else hi = mid - 1;
    }
    return lo;
}

}  // namespace dba
