// common.h - device-visible instance descriptor and helpers shared by all
// step kernels of the DBA library (gfx950 / wave64).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace dba {

// Address spaces for device code: the descriptor's data pointers are global
// memory (so accesses compile to global_load/global_store rather than flat
// instructions, which also count against the LDS wait counter), and the
// descriptor itself is read through the constant address space (scalar loads).
// In the host pass both are plain, so the layout is the same on both sides.
#if defined(__HIP_DEVICE_COMPILE__)
#define DBA_GLOBAL __attribute__((address_space(1)))
#define DBA_CONST __attribute__((address_space(4)))
#else
#define DBA_GLOBAL
#define DBA_CONST
#endif
typedef DBA_GLOBAL double gdouble;
typedef DBA_GLOBAL uint64_t gu64;
typedef DBA_GLOBAL uint8_t gu8;
typedef DBA_GLOBAL int32_t gi32;
typedef DBA_GLOBAL int64_t gi64;

constexpr int kMaxFields = 4;
constexpr int kMaxAux = 32;

// One simulation instance as the kernels see it. Lives in device memory
// (an array of these per merged group); everything the step needs is here.
struct Inst {
    gdouble* buf[2][kMaxFields];  // double-buffered prognostic fields (tile parity picks one)
    gdouble* aux[kMaxAux];        // model state kept for sleeping-neighbour semantics
    gu64* a_cur[2];               // awake bitmap of the step being executed, per set
    gu8* nflag[2];                // detector output per tile (Wakeup_Next), per set
    gu8* dF[2];                   // single-launch steps: detector bit per tile, by step parity
    gu8* dP[2];                   // ... buffer parity per tile, by step parity
    gi32* druns;                  // ... field-set updates per tile since the last canonicalisation
    gu64* par[2];                 // per-tile buffer parity per set: bit=1 -> buf[1] is current
    gdouble* tpart;               // conserved: per-tile change sums, [field][nt]
    gdouble* tspeed;              // KH: per-tile max wave speed, [0..nt) fresh (current U), [nt..2nt) stale
    gdouble* s_tile;              // conserved: per-tile epoch of the cumulative shift, [field][parity][nt]
    gdouble* scal;                // [1] dt of the next step, [2] time, [3] last dt, [8+f] cumulative shift S_f
    gi64* ctr;                    // [0] steps done, [1] active cell updates, [2] worklist length
    gi32* hist;                   // per-step active tile counts, [2*step + set]
    gi32* seg;                    // this instance's compacted worklist (ascending tiles)
    gu64* nbrec;                  // per worklist position (2D): the 3x3 neighbours' parity and awake bits
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
    double mb_cx, mb_cy, mb_cz, mb_lk, mb_omz;  // Model B: -1 + 1/r_x.., log(k_eq)/r_z, 1 - z_crit
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

// The descriptor as device code reads it (scalar loads).
using CInst = const DBA_CONST Inst;

// A per-tile counter read at the start of a workgroup and used at its end: a
// vector load, so it waits only where the value is used (a uniform address
// would otherwise become a scalar load that the wave waits for at once).
__device__ __forceinline__ int load_early(const gi32* p) {
    return __hip_atomic_load((int*)p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

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

__device__ __forceinline__ int bit_of(const gu64* bm, int t) {
    return (int)((bm[t >> 6] >> (t & 63)) & 1ull);
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

// Workgroup-wide OR of two predicates in one exchange: bit 0 = any(a), bit 1 = any(b).
// TAIL = false drops the closing barrier, for a last use of lds_scratch.
template <int NT, bool TAIL = true>
__device__ __forceinline__ int block_any2(int a, int b, int* lds_scratch) {
    constexpr int NW = NT / 64;
    const unsigned long long ba = __ballot(a), bb = __ballot(b);
    int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) lds_scratch[wid] = (ba != 0ull ? 1 : 0) | (bb != 0ull ? 2 : 0);
    __syncthreads();
    int r = 0;
#pragma unroll
    for (int w = 0; w < NW; w++) r |= lds_scratch[w];
    if (TAIL) __syncthreads();
    return r;
}

// Lane i gets v of lane i + N in its row of 16 lanes (a DPP row shift, no LDS
// traffic); lanes whose source leaves the row get 0.
template <int N>
__device__ __forceinline__ double dpp_row_down(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)b, 0x100 + N, 0xF, 0xF, true);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), 0x100 + N, 0xF, 0xF, true);
    return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}

// Lane 0 gets v summed over the wave by the shift tree v[i] += v[i + off],
// off = 32, 16, 8, 4, 2, 1: the two levels that cross 16-lane rows by lane
// permutes, the four inside a row by DPP row shifts. Lane 0 sees the operands
// of a plain shuffle tree in the same order, so the sum is the same double.
__device__ __forceinline__ double wave_sum_down(double v) {
    v += __shfl_down(v, 32, 64);
    v += __shfl_down(v, 16, 64);
    v += dpp_row_down<8>(v);
    v += dpp_row_down<4>(v);
    v += dpp_row_down<2>(v);
    v += dpp_row_down<1>(v);
    return v;
}

// Deterministic workgroup sum: per-wave shift reduction in a fixed order, then
// wave totals added in wave order by thread 0.
template <int NT>
__device__ __forceinline__ double block_sum(double v, double* lds_scratch) {
    constexpr int NW = NT / 64;
    v = wave_sum_down(v);
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

// The 3x3 block of tiles around tile t (2D), resolved once per tile into LDS:
// for each direction d = (dx+1) + 3 (dy+1), the tile that owns the halo cells
// on that side after the boundary map (wrap, or mirror back into t), its
// buffer parity per set, whether it runs this step, and its conserved shift.
struct NbTable {
    int tile[9];
    int par[2][9];
    int act[2][9];
    double shift[kMaxFields][9];
    double S;    // single-launch conserved steps: the cumulative shift of this step
    double tot;  // ... and the tile's change sum
};

// Owner of the halo cells of tile (ti, tj) (2D) in direction d = (dx+1) + 3 (dy+1):
// the tile holding the cells just outside the tile's edge on that side after
// the boundary map: the neighbour, wrapped when periodic, the tile itself when
// the boundary mirrors (zero-flux or wall).
__device__ __forceinline__ int nb_step1(int i, int di, int n, int bc) {
    if (di < 0) return i > 0 ? i - 1 : (bc == 0 ? n - 1 : i);
    if (di > 0) return i < n - 1 ? i + 1 : (bc == 0 ? 0 : i);
    return i;
}

__device__ __forceinline__ int nb_owner2d(CInst& I, int ti, int tj, int d) {
    return nb_step1(ti, d % 3 - 1, I.ntx, I.bc_x) + nb_step1(tj, d / 3 - 1, I.nty, I.bc_y == 0 ? 0 : 1) * I.ntx;
}

// Neighbour record written by the finalize kernel next to each worklist entry
// (2D models): bit s*18 + d = buffer parity, bit s*18 + 9 + d = awake, of the
// owner tile in direction d for set s, as the step reading it will see them;
// kRecValid marks a record whose nine owners all lie in the finalizing segment.
constexpr uint64_t kRecValid = 1ull << 36;

// Fill by threads 0..8; caller synchronises. With a valid record the parity and
// awake bits come from it; otherwise they are read from the bitmaps.
template <int TB>
__device__ __forceinline__ void nb_fill(NbTable& nb, CInst& I, int t, int x0, int y0, bool regular,
                                        int nsets, int ncons, uint64_t rec = 0) {
    (void)x0;
    (void)y0;
    const int d = threadIdx.x;
    if (d >= 9) return;
    const int tj0 = t / I.ntx;
    const int tt = nb_owner2d(I, t - tj0 * I.ntx, tj0, d);
    nb.tile[d] = tt;
    const bool have = (rec & kRecValid) != 0;
    for (int s = 0; s < nsets; s++) {
        if (have) {
            nb.par[s][d] = (int)((rec >> (s * 18 + d)) & 1ull);
            nb.act[s][d] = (int)((rec >> (s * 18 + 9 + d)) & 1ull);
        } else {
            nb.par[s][d] = bit_of(I.par[s], tt);
            nb.act[s][d] = regular ? 1 : bit_of(I.a_cur[s], tt);
        }
    }
    for (int f = 0; f < kMaxFields; f++)
        nb.shift[f][d] = f < ncons ? (I.scal[8 + f] - I.s_tile[((int64_t)f * 2 + nb.par[0][d]) * I.nt + tt]) : 0.0;
}

// Direction (0 lower, 1 the tile itself, 2 upper) along one axis of the owner of
// the mapped cell m, for tile i of n tiles of TB cells: the tile holding m is the
// tile, its upper neighbour (after the boundary map) or else its lower one. Used
// for tiles whose halo can reach past a partial last tile, where the owner no
// longer follows from the cell's offset in the tile's region.
__device__ __forceinline__ int owner_dir1(int m, int TB, int i, int n, int bc) {
    const int o = m / TB;
    if (o == i) return 1;
    return o == nb_step1(i, 1, n, bc) ? 2 : 0;
}

// Cells along one axis of tile i of tiles of TB cells on an axis of n cells:
// TB, except for a partial last tile.
__device__ __forceinline__ int tile_extent(int i, int TB, int n) { return min(TB, n - i * TB); }

// Direction index of a cell at local coordinate l (0..W-1) of a region whose
// interior [h, h+TB) is the tile.
template <int TB, int H>
__device__ __forceinline__ int nb_dir1(int l) {
    return l < H ? 0 : (l < TB + H ? 1 : 2);
}

template <int NT>
__device__ __forceinline__ double block_max(double v, double* lds_scratch) {
    constexpr int NW = NT / 64;
    // the shift tree of wave_sum_down with fmax (lane 0 sees the same operands:
    // none of the lanes it reads is a row shift's zero filler, so the result holds
    // for values of any sign; the same DPP shifts are pinned bit-exactly by the
    // field-sum test of dba_field_sum)
    v = fmax(v, __shfl_down(v, 32, 64));
    v = fmax(v, __shfl_down(v, 16, 64));
    v = fmax(v, dpp_row_down<8>(v));
    v = fmax(v, dpp_row_down<4>(v));
    v = fmax(v, dpp_row_down<2>(v));
    v = fmax(v, dpp_row_down<1>(v));
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

// A worklist segment: up to kSegTiles consecutive tiles of one instance. The
// finalize kernel compacts each segment into inst.seg[tile0 ...]; offsets is
// the exclusive prefix of the segment counts.
constexpr int kSegTiles = 1024;
struct Seg {
    int32_t inst;
    int32_t tile0;
    int32_t ntile;
    int32_t pad;
};

// Per-segment partial results of the finalize kernel, reduced per instance
// by the last workgroup in segment order (deterministic).
struct SegPart {
    unsigned long long c0, c1, cells;
    double csum[kMaxFields];
    double vmax;
};

// Per-instance segment range inside a group's segment table.
struct InstSegs {
    int32_t first, count;
};

// Item -> segment (offsets is the exclusive prefix, offsets[n] = total).
__device__ __forceinline__ int find_slot(const int32_t* offsets, int n, int item) {
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (offsets[mid] <= item) lo = mid; else hi = mid - 1;
    }
    return lo;
}

// The same segment for an item every lane of the wave holds: the offsets are
// non-decreasing, so the last segment k with offsets[k] <= item is the number of
// k in [1, n) with offsets[k] <= item. The wave reads 256 offsets per round trip
// (four per lane) and counts with ballots, where the binary search made
// log2(n) dependent loads (8 for a 256-segment list).
// With at most 257 segments (one round) the segment's offset is taken from the
// lane that loaded it, so the item's position needs no further load of it.
__device__ __forceinline__ int find_slot_wave(const int32_t* offsets, int n, int item, int* off_s) {
    const int lane = threadIdx.x & 63;
    int cnt = 0, o[4] = {0, 0, 0, 0};
    for (int k0 = 1; k0 < n; k0 += 256) {
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const int k = k0 + j * 64 + lane;
            o[j] = offsets[k < n ? k : 0];
        }
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const int k = k0 + j * 64 + lane;
            cnt += __popcll(__ballot(k < n && o[j] <= item));
        }
    }
    if (cnt == 0) {
        *off_s = 0;  // offsets[0]
    } else if (n <= 257) {
        const int i = cnt - 1, j = i >> 6;
        const int v = j == 0 ? o[0] : (j == 1 ? o[1] : (j == 2 ? o[2] : o[3]));
        *off_s = __builtin_amdgcn_readlane(v, i & 63);
    } else {
        *off_s = offsets[cnt];
    }
    return cnt;
}

// Current value of a field at cell i: the buffer the tile parity selects,
// minus the conserved shift accumulated since the tile last ran.
__device__ __forceinline__ double current_value(CInst& I, int field, int64_t i) {
    const int x = (int)(i % I.nx);
    const int y = (int)((i / I.nx) % I.ny);
    const int z = (int)(i / ((int64_t)I.nx * I.ny));
    const int t = x / I.tx + (y / I.ty) * I.ntx + (z / I.tz) * I.ntx * I.nty;
    const int set = I.field_set[field];
    const int p = bit_of(I.par[set], t);
    double v = I.buf[p][field][i];
    if (field < I.ncons) v -= (I.scal[8 + field] - I.s_tile[((int64_t)field * 2 + p) * I.nt + t]);
    return v;
}

// What a step kernel iterates over: the group's merged worklist. Groups of a
// single instance with a single segment also pass that instance's list and
// tile count, so the list entry, the list length and the descriptor are
// fetched in one round of loads instead of a dependent chain.
struct WorkList {
    const Inst* insts;
    const Seg* segs;
    const int32_t* offsets;   // exclusive prefix of segment counts, [nseg] = total
    int nseg;
    const int32_t* solo_seg;  // single-instance single-segment groups: its list, else null
    const uint64_t* solo_rec; // ... its neighbour records
    int solo_nt;              // ... and its tile count (0 otherwise)
    int one_inst;             // one instance: segment s holds its tiles from s kSegTiles on
};

// Item -> (instance, tile, neighbour record); false once the item is past the
// list. `total` is offsets[nseg], loaded by the caller once per workgroup.
// 3D models pass want_rec = false (their records are not written).
__device__ __forceinline__ bool next_item(const WorkList& w, int item, int total, CInst** I, int* t,
                                          uint64_t* rec, bool want_rec = true) {
    if (w.solo_nt > 0) {
        if (item >= w.solo_nt) return false;
        // in bounds; loaded before the length check so all three loads overlap
        const int tt = w.solo_seg[item];
        const uint64_t r = want_rec ? w.solo_rec[item] : 0ull;
        if (item >= total) return false;
        *I = (CInst*)w.insts;
        *t = __builtin_amdgcn_readfirstlane(tt);
        *rec = ((uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)r)) |
               ((uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(r >> 32)) << 32);
        return true;
    }
    if (item >= total) return false;
    int s = 0, off = 0;
    if (w.nseg > 1) s = find_slot_wave(w.offsets, w.nseg, item, &off);
    int inst = 0, tile0 = s * kSegTiles;
    if (!w.one_inst) {
        const Seg sg = w.segs[s];
        inst = sg.inst;
        tile0 = sg.tile0;
    }
    *I = (CInst*)w.insts + __builtin_amdgcn_readfirstlane(inst);
    const int pos = __builtin_amdgcn_readfirstlane(tile0 + (item - off));
    *t = __builtin_amdgcn_readfirstlane((*I)->seg[pos]);
    const uint64_t r = want_rec ? (*I)->nbrec[pos] : 0ull;
    *rec = ((uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)r)) |
           ((uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(r >> 32)) << 32);
    return true;
}

// Grid-stride walk over the worklist with an XCD-aware deal: workgroups are
// issued to the 8 XCDs in turn (b % 8), so each round of G items is split into
// 8 contiguous ranges, range x going to the workgroups of XCD x. A round with
// fewer than G items (a sparse list) is still spread over all 8 XCDs. For a
// single-instance list, the first item of a full round is read before the list
// length arrives (it is the same item either way), so both loads overlap.
struct WorkCursor {
    int base;       // first item of the next round
    int spec_item;  // item whose entry was read ahead (-1: none)
    int spec_t;
    uint64_t spec_rec;
};

__device__ __forceinline__ int xcd_deal(int b, int G, int base, int cnt) {
    if ((G & 7) != 0) return b < cnt ? base + b : -1;
    const int per = (cnt + 7) >> 3, slot = b >> 3, x = b & 7;
    if (slot >= per) return -1;
    const int item = base + x * per + slot;
    return item < base + cnt ? item : -1;
}

__device__ __forceinline__ void work_begin(const WorkList& w, WorkCursor& c, bool want_rec) {
    c.base = 0;
    c.spec_item = -1;
    if (w.solo_nt > 0) {
        const int G = gridDim.x;
        const int item = xcd_deal(blockIdx.x, G, 0, G);  // the item of a full first round
        if (item >= 0 && item < w.solo_nt) {
            c.spec_item = item;
            c.spec_t = w.solo_seg[item];
            c.spec_rec = want_rec ? w.solo_rec[item] : 0ull;
        }
    }
}

__device__ __forceinline__ bool work_next(const WorkList& w, WorkCursor& c, int total, CInst** I, int* t,
                                          uint64_t* rec, bool want_rec = true) {
    const int G = gridDim.x;
    while (c.base < total) {
        const int item = xcd_deal(blockIdx.x, G, c.base, min(G, total - c.base));
        c.base += G;
        if (item < 0) continue;
        if (item == c.spec_item) {
            *I = (CInst*)w.insts;
            *t = __builtin_amdgcn_readfirstlane(c.spec_t);
            *rec = ((uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)c.spec_rec)) |
                   ((uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(c.spec_rec >> 32)) << 32);
            c.spec_item = -1;
            return true;
        }
        return next_item(w, item, total, I, t, rec, want_rec);
    }
    return false;
}

// The blocks one workgroup visits in 64 consecutive grid-stride rounds of the
// deal work_next makes, resolved at once (lane j: round round0 + j), so a kernel
// can read blocks ahead without a dependent worklist load per block.
struct RoundList {
    uint64_t valid;  // bit j: this workgroup has a block in round round0 + j
    int inst;        // lane j: that block's instance
    int t;           // lane j: that block
};

__device__ __forceinline__ RoundList round_list(const WorkList& w, int total, int round0) {
    const int G = gridDim.x, lane = threadIdx.x & 63;
    const int64_t base = (int64_t)(round0 + lane) * G;
    int item = -1;
    if (base < total) item = xcd_deal(blockIdx.x, G, (int)base, (int)min((int64_t)G, (int64_t)total - base));
    RoundList L;
    L.inst = 0;
    L.t = 0;
    if (item >= 0) {
        if (w.solo_nt > 0) {
            L.t = w.solo_seg[item];
        } else {
            const int s = w.nseg == 1 ? 0 : find_slot(w.offsets, w.nseg, item);
            const Seg sg = w.segs[s];
            L.inst = sg.inst;
            L.t = w.insts[sg.inst].seg[sg.tile0 + (item - w.offsets[s])];
        }
    }
    L.valid = __ballot(item >= 0);
    return L;
}

// Pops the earliest remaining round of `todo` (uniform) and returns its block.
__device__ __forceinline__ void round_take(const WorkList& w, const RoundList& L, uint64_t& todo, CInst** I,
                                           int* t) {
    const int r = __builtin_ctzll(todo);
    todo &= todo - 1;
    *I = (CInst*)w.insts + __builtin_amdgcn_readlane(L.inst, r);
    *t = __builtin_amdgcn_readlane(L.t, r);
}

}  // namespace dba
