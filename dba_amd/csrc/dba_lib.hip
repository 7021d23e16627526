// dba_lib.hip - C ABI of the MI355X DBA time step (libdba_hip.so).
//
// Host orchestration of the gfx950 kernels: instances own their device state,
// groups merge instances of one model into one worklist, every step is
//   step kernel (fused stencil + detector byte per tile)
//   -> finalize kernel (parity flip, history, conserved shift / CFL dt,
//      compaction of the next worklist),
// on one HIP stream, optionally replayed from a captured hipGraph.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "../../include/dba/dba.h"
#include "common.h"
#include "k_finalize.h"
#include "k_ch.h"
#include "k_k93.h"
#include "k_kh.h"
#include "k_ref3d.h"
#include "k_diag.h"
#include "k_droplets.h"

using namespace dba;

// Device scratch of dba_droplets_noalloc: one array per GPU, held from module
// load until the process ends. Host code reaches it only through the address
// hipGetSymbolAddress returns for the current device (dba_ctx::scratch); the
// name itself is a host-side stand-in, not a device address.
constexpr int64_t kScratchMaxCells = int64_t(1) << 22;  // 2048^2 (161^3) cells
constexpr size_t kScratchTail = 64 * 1024;               // block partials, counters
constexpr size_t kScratchBytes = 13 * (size_t)kScratchMaxCells + kScratchTail;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

thread_local std::string g_error;

dba_status fail(dba_status code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_error = buf;
    return code;
}

#define HIP_TRY(expr)                                                                  \
    do {                                                                               \
        hipError_t e_ = (expr);                                                        \
        if (e_ != hipSuccess)                                                          \
            return fail(DBA_ERR_HIP, "%s failed: %s", #expr, hipGetErrorString(e_));   \
    } while (0)

#define LAUNCH_CHECK()                                                                 \
    do {                                                                               \
        hipError_t e_ = hipGetLastError();                                             \
        if (e_ != hipSuccess)                                                          \
            return fail(DBA_ERR_LAUNCH, "kernel launch failed: %s", hipGetErrorString(e_)); \
    } while (0)

int nfields_of(int model) {
    switch (model) {
        case DBA_MODEL_CH: return 1;
        case DBA_MODEL_DENDRITE_K93: return 2;
        case DBA_MODEL_KH: return 4;
        case DBA_MODEL_DENDRITE_K94: return 2;
        case DBA_MODEL_MODELB: return 4;
        default: return 0;
    }
}

int nsets_of(int model) {
    return (model == DBA_MODEL_DENDRITE_K93 || model == DBA_MODEL_DENDRITE_K94) ? 2 : 1;
}

bool is_pow2_tile(int t) { return t == 8 || t == 16 || t == 32; }

struct Instance {
    int id = 0;
    dba_params p{};
    Inst h{};                    // host copy of the device descriptor
    std::vector<void*> allocs;   // device allocations owned by this instance
    int64_t footprint = 0;
    int64_t ncell = 0;
    bool needs_prime = false;    // KH: per-tile wave speeds must be recomputed before the next step
    bool stepped = false;        // work for this instance has been enqueued on the compute stream
    Inst* d_self = nullptr;      // single-instance descriptor (downloads, checks)
};

struct Group {
    std::vector<int> ids;
    int model = 0, tile = 0, tile_z = 1, threads = 0;
    int dense_threads = 0;        // workgroup size of the single-launch step kernel
    bool conserved = false;       // one conserved instance (single-launch steps carry its mean shift)
    bool partial = false;         // an instance with partial tiles: the kernels' PART instances
    int march = 0;                // column-march single-launch steps: tiles per strip (0: off)
    bool spec = false;            // K93 single-launch steps with speculative staging
    bool kh_p2 = false;           // KH: dx, dy and the cell area of every instance are powers of two
    int sb_from = 0;              // Model B: lists of at least this many blocks step on superblocks (0: never)
    bool nine = true;             // CH 9-point stencil (the march kernel is instanced on it)
    Inst* d_insts = nullptr;
    Seg* d_segs = nullptr;
    InstSegs* d_isegs = nullptr;
    SegPart* d_parts = nullptr;
    int32_t* d_counts = nullptr;
    int32_t* d_offsets = nullptr;
    int32_t* d_dmap = nullptr;    // single-launch steps of several instances or segments: instance << 24 | tile
    int nseg = 0;
    int64_t total_tiles = 0;
    int grid = 0;
    uint64_t last_use = 0;
    int32_t* solo_seg = nullptr;  // one instance in one segment: its worklist ...
    uint64_t* solo_rec = nullptr; // ... its neighbour records
    int solo_nt = 0;              // ... and tile count
};

struct GraphKey {
    std::vector<int> ids;
    int chunk;
    int single_launch;  // the mode the graph was captured in
    bool operator<(const GraphKey& o) const {
        if (chunk != o.chunk) return chunk < o.chunk;
        if (single_launch != o.single_launch) return single_launch < o.single_launch;
        return ids < o.ids;
    }
};

}  // namespace

struct CachedGraph {
    hipGraphExec_t exec = nullptr;
    uint64_t last_use = 0;
};

struct dba_ctx {
    int device = 0;
    hipStream_t stream = nullptr;  // compute stream: steps, finalize, downloads
    hipStream_t setup = nullptr;   // instance and group set-up, uploads of never-stepped instances
    int next_id = 1;
    uint64_t clock = 0;            // LRU clock of the group and graph caches
    std::map<int, Instance*> inst;
    std::map<std::vector<int>, Group*> groups;
    std::map<GraphKey, CachedGraph> graphs;
    double* staging = nullptr;
    size_t staging_n = 0;
    int* d_flag = nullptr;
    double* d_diag = nullptr;      // diagnostics: block partials, results
    int32_t* d_cols = nullptr;     // diagnostics: growth-front columns
    bool timing = false;
    int single_launch = 1;         // single-launch steps: 0 off, 1 one-instance groups, 2 also batches
    std::vector<std::pair<hipEvent_t, hipEvent_t>> ev_step, ev_fin;  // timed intervals (events may be shared)
    std::vector<hipEvent_t> ev_used;  // every event handed out since the timing was last reset
    std::vector<hipEvent_t> ev_pool;
    hipEvent_t markers[8] = {};
    unsigned char* scratch = nullptr;  // this device's address of the scratch array
};

namespace {

// Derived constants: the same C expressions the oracle evaluates.
void derive_constants(const dba_params& p, Inst& h) {
    h.invdx2 = 1.0 / (p.dx * p.dx);
    h.inv6dx2 = 1.0 / (6.0 * p.dx * p.dx);
    h.inv2dx = 1.0 / (2.0 * p.dx);
    h.dt_tau = p.tau != 0.0 ? p.dt / p.tau : 0.0;
    h.alpha_pi = p.alpha / M_PI;
    double jf = (double)p.aniso_j;
    h.cos_jt0 = std::cos(jf * p.theta0);
    h.sin_jt0 = std::sin(jf * p.theta0);
    h.k_epsp = -(p.eps_bar * jf * p.delta);
    if (p.r_x != 0.0 && p.r_y != 0.0 && p.r_z != 0.0) {  // Model B (validated there)
        h.mb_cx = -1.0 + 1.0 / p.r_x;
        h.mb_cy = -1.0 + 1.0 / p.r_y;
        h.mb_cz = -1.0 + 1.0 / p.r_z;
        h.mb_lk = std::log(p.k_eq) / p.r_z;
    }
    h.mb_omz = 1.0 - p.z_crit;
}

dba_status validate(const dba_params& p) {
    if (nfields_of(p.model) == 0) return fail(DBA_ERR_INVALID, "model %d is not supported", p.model);
    if (p.mode < 0 || p.mode > 2) return fail(DBA_ERR_INVALID, "mode %d invalid", p.mode);
    if (p.nx <= 0 || p.ny <= 0 || p.nz <= 0) return fail(DBA_ERR_INVALID, "grid %dx%dx%d invalid", p.nx, p.ny, p.nz);
    if (p.model == DBA_MODEL_CH || p.model == DBA_MODEL_DENDRITE_K93) {
        if (p.nz != 1) return fail(DBA_ERR_INVALID, "model %d is 2D (nz must be 1)", p.model);
        if (p.tile_x != p.tile_y || !is_pow2_tile(p.tile_x))
            return fail(DBA_ERR_INVALID, "tile must be square 8, 16 or 32 (got %dx%d)", p.tile_x, p.tile_y);
        // a partial last tile needs two cells or more: its neighbours' halos and its
        // own ring stay within the 3x3 tiles around each tile
        if (p.nx % p.tile_x == 1 || p.ny % p.tile_y == 1)
            return fail(DBA_ERR_INVALID, "grid %dx%d leaves a partial tile one cell wide (tile %d)", p.nx, p.ny,
                        p.tile_x);
        if (p.nx < p.tile_x * 1 || p.ny < p.tile_y) return fail(DBA_ERR_INVALID, "grid smaller than a tile");
    }
    if (p.model == DBA_MODEL_KH) {
        if (p.nz != 1) return fail(DBA_ERR_INVALID, "KH is 2D (nz must be 1)");
        if (p.tile_x != p.tile_y || (p.tile_x != 8 && p.tile_x != 16))
            return fail(DBA_ERR_INVALID, "KH tile must be 8x8 or 16x16");
        if (p.nx % p.tile_x || p.ny % p.tile_y)
            return fail(DBA_ERR_INVALID, "grid %dx%d must be a multiple of the tile %d", p.nx, p.ny, p.tile_x);
        if (p.bc_x != DBA_BC_PERIODIC) return fail(DBA_ERR_INVALID, "KH needs periodic x");
        if (p.bc_y != DBA_BC_PERIODIC && p.bc_y != DBA_BC_WALL) return fail(DBA_ERR_INVALID, "KH y: periodic or wall");
        if (p.dt_mode == DBA_DT_CFL && p.mode == DBA_MODE_CONSERVED)
            return fail(DBA_ERR_INVALID, "conserved KH uses a fixed dt");
        if (p.dt_mode == DBA_DT_CFL && !(p.cfl > 0.0)) return fail(DBA_ERR_INVALID, "cfl must be positive");
        if (!(p.gas_gamma > 1.0)) return fail(DBA_ERR_INVALID, "gas_gamma must exceed 1");
        if (p.dy != p.dx) return fail(DBA_ERR_INVALID, "KH needs dy == dx");
    }
    if (p.model == DBA_MODEL_DENDRITE_K94 || p.model == DBA_MODEL_MODELB) {
        const int tz = p.tile_z > 0 ? p.tile_z : 1;
        if (p.tile_x != 4 || p.tile_y != 4 || (tz != 4 && tz != 1) || (tz == 1) != (p.nz == 1))
            return fail(DBA_ERR_INVALID, "reference 3D models use 4x4x4 blocks (4x4x1 when nz == 1)");
        if (p.nx % 4 || p.ny % 4 || p.nz % tz)
            return fail(DBA_ERR_INVALID, "grid %dx%dx%d must be a multiple of the block", p.nx, p.ny, p.nz);
        if (p.bc_x != DBA_BC_PERIODIC || p.bc_y != DBA_BC_PERIODIC || p.bc_z != DBA_BC_PERIODIC)
            return fail(DBA_ERR_INVALID, "reference 3D models are periodic");
        if (p.model == DBA_MODEL_DENDRITE_K94 && p.mode == DBA_MODE_CONSERVED)
            return fail(DBA_ERR_INVALID, "conserved DBA is defined for Model B and KH only");
        if (p.model == DBA_MODEL_MODELB && (p.r_x == 0.0 || p.r_y == 0.0 || p.r_z == 0.0))
            return fail(DBA_ERR_INVALID, "Model B needs nonzero rx, ry, rz");
    }
    if (p.model == DBA_MODEL_CH && p.stencil != 5 && p.stencil != 9)
        return fail(DBA_ERR_INVALID, "CH stencil must be 5 or 9");
    if (p.model == DBA_MODEL_CH && p.mode == DBA_MODE_CONSERVED && p.bc_x != DBA_BC_PERIODIC)
        return fail(DBA_ERR_INVALID, "conserved DBA needs a periodic grid");
    if (p.model == DBA_MODEL_DENDRITE_K93 && (p.aniso_j < 1 || p.aniso_j > 16))
        return fail(DBA_ERR_INVALID, "aniso_j must be in [1,16]");
    if (p.model == DBA_MODEL_DENDRITE_K93 && p.mode == DBA_MODE_CONSERVED)
        return fail(DBA_ERR_INVALID, "conserved DBA is defined for Model B and KH only");
    const bool wide = p.threads == 512 || p.threads == 1024;
    if (p.threads != 0 && p.threads != 64 && p.threads != 128 && p.threads != 256 && !wide)
        return fail(DBA_ERR_INVALID, "threads must be 0, 64, 128, 256, 512 or 1024");
    if (wide && p.model != DBA_MODEL_CH && p.model != DBA_MODEL_DENDRITE_K93)
        return fail(DBA_ERR_INVALID, "512 and 1024 threads are for the CH and Kobayashi-93 models");
    if (p.dx <= 0.0 || p.dt <= 0.0) return fail(DBA_ERR_INVALID, "dx and dt must be positive");
    return DBA_OK;
}

struct Layout {
    int ntx, nty, ntz, nt, nwords;
    int64_t n;
    int nf, nsets, naux;
    int64_t aux_len;  // doubles per aux array
};

Layout layout_of(const dba_params& p) {
    Layout L{};
    int tz = p.tile_z > 0 ? p.tile_z : 1;
    L.ntx = (p.nx + p.tile_x - 1) / p.tile_x;
    L.nty = (p.ny + p.tile_y - 1) / p.tile_y;
    L.ntz = (p.nz + tz - 1) / tz;
    L.nt = L.ntx * L.nty * L.ntz;
    L.nwords = (L.nt + 63) / 64;
    L.n = (int64_t)p.nx * p.ny * p.nz;
    L.nf = nfields_of(p.model);
    L.nsets = nsets_of(p.model);
    L.naux = 0;
    L.aux_len = L.n;
    if (p.model == DBA_MODEL_MODELB) L.naux = 3;  // chemical potentials (every cell, every mode)
    if (p.mode != DBA_MODE_REGULAR) {
        if (p.model == DBA_MODEL_CH) {             // mu ring: one record per tile
            L.naux = 1;
            L.aux_len = (int64_t)L.nt * (4 * p.tile_x - 4);
        }
        if (p.model == DBA_MODEL_KH) {             // edge state of sleeping tiles: one record per tile
            L.naux = 1;
            L.aux_len = (int64_t)L.nt * (p.tile_x == 8 ? KhEdge<8>::R : KhEdge<16>::R);
        }
    }
    return L;
}

int64_t footprint_of(const dba_params& p) {
    Layout L = layout_of(p);
    int64_t b = 0;
    b += 2 * (int64_t)L.nf * L.n * 8;        // double-buffered fields
    b += (int64_t)L.naux * L.aux_len * 8;    // stored neighbour state
    b += 4 * (int64_t)L.nwords * 8 + 2 * (int64_t)L.nt;  // bitmaps + detector bytes
    b += (int64_t)kMaxFields * L.nt * 8;     // tpart
    b += 2 * (int64_t)L.nt * 8;              // tspeed
    b += (int64_t)kMaxFields * 2 * L.nt * 8; // s_tile
    b += 16 * 8 + 4 * 8;                     // scal + ctr
    b += (int64_t)2 * std::max(0, p.history_cap) * 4;
    b += (int64_t)L.nt * 4;                  // seg
    b += (int64_t)L.nt * 8;                  // nbrec
    b += sizeof(Inst);
    return b;
}

template <typename T>
dba_status dalloc(Instance* in, T** out, size_t count, hipStream_t setup) {
    void* ptr = nullptr;
    size_t bytes = std::max<size_t>(count * sizeof(T), 8);
    hipError_t e = hipMalloc(&ptr, bytes);
    if (e != hipSuccess) return fail(DBA_ERR_NOMEM, "hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
    e = hipMemsetAsync(ptr, 0, bytes, setup);
    if (e != hipSuccess) return fail(DBA_ERR_HIP, "hipMemset failed: %s", hipGetErrorString(e));
    in->allocs.push_back(ptr);
    *out = (T*)ptr;
    return DBA_OK;
}

void free_instance(Instance* in) {
    for (void* p : in->allocs) (void)hipFree(p);
    in->allocs.clear();
}

void free_group(Group* g) {
    (void)hipFree(g->d_insts);
    (void)hipFree(g->d_segs);
    (void)hipFree(g->d_isegs);
    (void)hipFree(g->d_parts);
    (void)hipFree(g->d_counts);
    (void)hipFree(g->d_offsets);
    (void)hipFree(g->d_dmap);
    delete g;
}

bool subset(const std::vector<int>& a, const std::vector<int>& b) {  // every id of a is in b
    for (int x : a)
        if (std::find(b.begin(), b.end(), x) == b.end()) return false;
    return true;
}

// Drops every captured graph that may launch group g (its ids are a subset of
// the graph's instance set).
void drop_graphs_using(dba_ctx* ctx, const std::vector<int>& gids) {
    for (auto it = ctx->graphs.begin(); it != ctx->graphs.end();) {
        if (subset(gids, it->first.ids)) {
            (void)hipGraphExecDestroy(it->second.exec);
            it = ctx->graphs.erase(it);
        } else {
            ++it;
        }
    }
}

void invalidate_groups(dba_ctx* ctx, int id) {
    for (auto it = ctx->groups.begin(); it != ctx->groups.end();) {
        if (std::find(it->first.begin(), it->first.end(), id) != it->first.end()) {
            free_group(it->second);
            it = ctx->groups.erase(it);
        } else {
            ++it;
        }
    }
    drop_graphs_using(ctx, {id});
}

// Bounded caches: a serving loop steps a changing set of instances, and every
// distinct set has its own groups and graphs. Entries not used by the current
// call (last_use < clock) are evicted least recently used first.
constexpr size_t kMaxGroups = 64;
constexpr size_t kMaxGraphs = 32;

void evict_groups(dba_ctx* ctx) {
    while (ctx->groups.size() >= kMaxGroups) {
        auto victim = ctx->groups.end();
        for (auto it = ctx->groups.begin(); it != ctx->groups.end(); ++it)
            if (it->second->last_use < ctx->clock &&
                (victim == ctx->groups.end() || it->second->last_use < victim->second->last_use))
                victim = it;
        if (victim == ctx->groups.end()) return;
        (void)hipStreamSynchronize(ctx->stream);
        drop_graphs_using(ctx, victim->first);
        free_group(victim->second);
        ctx->groups.erase(victim);
    }
}

void evict_graphs(dba_ctx* ctx) {
    while (ctx->graphs.size() >= kMaxGraphs) {
        auto victim = ctx->graphs.end();
        for (auto it = ctx->graphs.begin(); it != ctx->graphs.end(); ++it)
            if (it->second.last_use < ctx->clock &&
                (victim == ctx->graphs.end() || it->second.last_use < victim->second.last_use))
                victim = it;
        if (victim == ctx->graphs.end()) return;
        (void)hipStreamSynchronize(ctx->stream);
        (void)hipGraphExecDestroy(victim->second.exec);
        ctx->graphs.erase(victim);
    }
}

// Workgroup size per (model, tile) measured best on MI355X (CH and Kobayashi-93
// grids of 1024^2 to 4096^2); the workgroup size never changes results.
int default_threads(int model, int tile) {
    if (tile <= 8) return 64;
    if (tile == 16) return (model == DBA_MODEL_CH || model == DBA_MODEL_DENDRITE_K93) ? 128 : 256;
    return model == DBA_MODEL_DENDRITE_K93 ? 512 : 256;
}

// Single-launch steps of K93 16x16 tiles: every tile takes a workgroup, so a
// small grid gains from more threads per tile (a shorter FP64 chain per tile)
// and a large, mostly sleeping one from narrower workgroups (more of them per
// CU). Measured on MI355X: 512^2 (1024 tiles) 512 threads 8.0 us vs 128 threads
// 10.0 us; 1024^2 256 threads 9.9 vs 11.8 us; 2048^2 128 threads best.
int default_dense_threads(int model, int tile, int64_t tiles) {
    if (model == DBA_MODEL_DENDRITE_K93 && tile == 16) return tiles <= 1024 ? 512 : (tiles <= 4096 ? 256 : 128);
    return default_threads(model, tile);
}

// KH: wave speeds of every tile from the current state, and the first CFL dt.
dba_status kh_prime(dba_ctx* ctx, Instance* in) {
    if (in->p.tile_x == 8)
        hipLaunchKernelGGL((kh_prime_kernel<8>), dim3(std::min(in->h.nt, 4096)), dim3(256), 0, ctx->stream,
                           in->d_self, 0);
    else
        hipLaunchKernelGGL((kh_prime_kernel<16>), dim3(std::min(in->h.nt, 4096)), dim3(256), 0, ctx->stream,
                           in->d_self, 0);
    LAUNCH_CHECK();
    hipLaunchKernelGGL(kh_dt_kernel, dim3(1), dim3(256), 0, ctx->stream, in->d_self, 0);
    LAUNCH_CHECK();
    in->needs_prime = false;
    return DBA_OK;
}

dba_status get_group(dba_ctx* ctx, const std::vector<int>& ids, Group** out) {
    auto it = ctx->groups.find(ids);
    if (it != ctx->groups.end()) {
        it->second->last_use = ctx->clock;
        *out = it->second;
        return DBA_OK;
    }
    evict_groups(ctx);
    Group* g = new Group();
    g->last_use = ctx->clock;
    g->ids = ids;
    std::vector<Inst> hs;
    std::vector<Seg> segs;
    std::vector<InstSegs> isegs;
    for (size_t k = 0; k < ids.size(); k++) {
        Instance* in = ctx->inst[ids[k]];
        hs.push_back(in->h);
        g->total_tiles += in->h.nt;
        InstSegs r{(int32_t)segs.size(), 0};
        for (int t0 = 0; t0 < in->h.nt; t0 += kSegTiles) {
            segs.push_back(Seg{(int32_t)k, t0, std::min(kSegTiles, in->h.nt - t0), 0});
            r.count++;
        }
        isegs.push_back(r);
    }
    g->nseg = (int)segs.size();
    Instance* first = ctx->inst[ids[0]];
    if (ids.size() == 1 && g->nseg == 1) {
        g->solo_seg = first->h.seg;
        g->solo_rec = first->h.nbrec;
        g->solo_nt = first->h.nt;
    }
    g->model = first->p.model;
    g->tile = first->p.tile_x;
    g->tile_z = first->p.tile_z;
    g->threads = first->p.threads ? first->p.threads : default_threads(g->model, g->tile);
    g->dense_threads = first->p.threads ? first->p.threads : default_dense_threads(g->model, g->tile, g->total_tiles);
    g->conserved = ids.size() == 1 && first->p.mode == DBA_MODE_CONSERVED;
    for (int id : ids)
        if (!ctx->inst[id]->h.full_tiles) g->partial = true;
    // Column-march single-launch steps (ch_march_kernel): one large CH instance of
    // full 32 x 32 tiles, periodic both ways, not conserved; K tiles per strip,
    // the most that keep >= 1024 strips (measured: 2048^2 K = 4 18.4 us against
    // 20.8 at K = 2 and 22.5 at K = 8; 4096^2 K = 16 59.7 us, K = 8 60.2 us, K = 4
    // 63.1 us; 8192^2 K = 32 248.9 us against 253.0 at K = 8)
    g->nine = first->p.stencil != 5;
    if (ids.size() == 1 && g->model == DBA_MODEL_CH && g->tile == 32 && g->dense_threads == 256 &&
        first->p.mode != DBA_MODE_CONSERVED && first->h.full_tiles && first->p.bc_x == DBA_BC_PERIODIC &&
        first->p.bc_y == DBA_BC_PERIODIC && first->h.nt >= kMarchMinTiles) {
        for (int k : {32, 16, 8, 4, 2})
            if (first->h.nty % k == 0 && first->h.nt / k >= 1024) {
                g->march = k;
                break;
            }
    }
    // Speculative staging in single-launch K93 steps (k93_dense_kernel SPEC):
    // one instance of up to 1024 full 16 x 16 tiles at 512 threads (measured on
    // 512^2: 7.2 -> 6.7 us per step with the run-counter prefetch); larger grids
    // would pay the rings of both buffers of every sleeping tile in HBM bytes
    g->spec = ids.size() == 1 && g->model == DBA_MODEL_DENDRITE_K93 && g->tile == 16 && g->dense_threads == 512 &&
              first->h.full_tiles && first->h.nt <= 1024;
    // KH divides by dx, dy and the cell area; when all three are powers of two
    // (a power-of-two grid on the unit square) the step multiplies by their exact
    // reciprocals instead, which gives the same doubles (kdiv in k_kh.h)
    if (g->model == DBA_MODEL_KH) {
        auto pow2 = [](double d) {
            int e = 0;
            return d > 0.0 && std::isfinite(d) && std::frexp(d, &e) == 0.5;
        };
        g->kh_p2 = true;
        for (int id : ids) {
            const Instance* in = ctx->inst.at(id);
            if (!(pow2(in->h.dx) && pow2(in->h.dy) && pow2(in->h.vol))) g->kh_p2 = false;
        }
    }
    // Model B: 65536 workgroups (4 blocks each on a fully awake 256^3 field)
    // measured faster than 16384 or 3072 with longer pipelines per wave
    int64_t cap = g->model == DBA_MODEL_MODELB ? 65536 : (g->model == DBA_MODEL_DENDRITE_K94 ? 16384 : 4096);
    g->grid = (int)std::max<int64_t>(1, std::min<int64_t>(g->total_tiles, cap));
    // Model B superblocks (modelb_mu_sb_kernel): one 3D instance of full 4x4x4
    // blocks, an even number of blocks in x and y, no conserved shift; steps with
    // at least a quarter of the blocks awake. DBA_MODELB_SUPERBLOCKS=0 turns them
    // off and =1 uses them for every non-empty list (both give the same doubles).
    if (g->model == DBA_MODEL_MODELB && ids.size() == 1 && g->tile == 4 && g->tile_z == 4) {
        const Inst& h = first->h;
        if (first->p.mode != DBA_MODE_CONSERVED && h.full_tiles && h.nz > 1 && h.ntx % 2 == 0 &&
            h.nty % 2 == 0) {
            const char* env = std::getenv("DBA_MODELB_SUPERBLOCKS");
            const std::string e = env ? env : "";
            g->sb_from = e == "0" ? 0 : (e == "1" ? 1 : (int)((h.nt + 3) / 4));
            // the block kernels then only take lists shorter than sb_from
            if (g->sb_from > 0)
                g->grid = (int)std::max<int64_t>(1, std::min<int64_t>(g->grid, (g->sb_from + 7) / 8));
        }
    }
    size_t n = ids.size();
    HIP_TRY(hipMalloc(&g->d_insts, sizeof(Inst) * n));
    HIP_TRY(hipMemcpyAsync(g->d_insts, hs.data(), sizeof(Inst) * n, hipMemcpyHostToDevice, ctx->setup));
    size_t ns = segs.size();
    HIP_TRY(hipMalloc(&g->d_segs, sizeof(Seg) * ns));
    HIP_TRY(hipMemcpyAsync(g->d_segs, segs.data(), sizeof(Seg) * ns, hipMemcpyHostToDevice, ctx->setup));
    HIP_TRY(hipMalloc(&g->d_isegs, sizeof(InstSegs) * n));
    HIP_TRY(hipMemcpyAsync(g->d_isegs, isegs.data(), sizeof(InstSegs) * n, hipMemcpyHostToDevice, ctx->setup));
    HIP_TRY(hipMalloc(&g->d_parts, sizeof(SegPart) * ns));
    HIP_TRY(hipMemsetAsync(g->d_parts, 0, sizeof(SegPart) * ns, ctx->setup));
    HIP_TRY(hipMalloc(&g->d_counts, sizeof(int32_t) * ns));
    HIP_TRY(hipMalloc(&g->d_offsets, sizeof(int32_t) * (ns + 1)));
    HIP_TRY(hipMemsetAsync(g->d_offsets, 0, sizeof(int32_t) * (ns + 1), ctx->setup));
    std::vector<int32_t> dmap;
    if ((n > 1 || ns > 1) && (g->model == DBA_MODEL_CH || g->model == DBA_MODEL_DENDRITE_K93)) {
        // dense index -> instance << 24 | tile (groups hold at most 255 instances)
        bool fits = true;
        for (size_t k = 0; k < n; k++) fits = fits && hs[k].nt < (1 << 24);
        if (fits) {
            dmap.reserve((size_t)g->total_tiles);
            for (size_t k = 0; k < n; k++)
                for (int t = 0; t < hs[k].nt; t++) dmap.push_back((int32_t)((uint32_t)k << 24 | (uint32_t)t));
            HIP_TRY(hipMalloc(&g->d_dmap, sizeof(int32_t) * dmap.size()));
            HIP_TRY(hipMemcpyAsync(g->d_dmap, dmap.data(), sizeof(int32_t) * dmap.size(), hipMemcpyHostToDevice,
                                   ctx->setup));
        }
    }
    // the compute stream is used only after this host-side wait
    HIP_TRY(hipStreamSynchronize(ctx->setup));
    ctx->groups[ids] = g;
    *out = g;
    return DBA_OK;
}

// ---------------------------------------------------------------- launches

WorkList wl_of(const Group* g) {
    return WorkList{g->d_insts, g->d_segs, g->d_offsets, g->nseg, g->solo_seg, g->solo_rec, g->solo_nt,
                    g->ids.size() == 1 ? 1 : 0};
}

template <int TB, int NT>
void launch_step_tb(const Group* g, hipStream_t s) {
    dim3 grid(g->grid), block(NT);
    switch (g->model) {
        case DBA_MODEL_CH: {
            // the strip layout needs NT <= TB*TB
            constexpr int NTC = NT <= TB * TB ? NT : TB * TB;
            if (g->partial)
                hipLaunchKernelGGL((ch_step_kernel<TB, NTC, true>), grid, dim3(NTC), 0, s, wl_of(g));
            else
                hipLaunchKernelGGL((ch_step_kernel<TB, NTC, false>), grid, dim3(NTC), 0, s, wl_of(g));
            break;
        }
        case DBA_MODEL_DENDRITE_K93:
            if (g->partial)
                hipLaunchKernelGGL((k93_step_kernel<TB, NT, true>), grid, block, 0, s, wl_of(g));
            else
                hipLaunchKernelGGL((k93_step_kernel<TB, NT, false>), grid, block, 0, s, wl_of(g));
            break;
        case DBA_MODEL_KH:
            if constexpr (TB <= 16 && NT <= 256) {
                if (g->kh_p2)
                    hipLaunchKernelGGL((kh_step_kernel<TB, NT, true>), grid, block, 0, s, wl_of(g));
                else
                    hipLaunchKernelGGL((kh_step_kernel<TB, NT, false>), grid, block, 0, s, wl_of(g));
            }
            break;
        default:
            break;
    }
}

template <int TB>
void launch_step_nt(const Group* g, hipStream_t s) {
    switch (g->threads) {
        case 64: launch_step_tb<TB, 64>(g, s); break;
        case 128: launch_step_tb<TB, 128>(g, s); break;
        case 512: launch_step_tb<TB, 512>(g, s); break;
        case 1024: launch_step_tb<TB, 1024>(g, s); break;
        default: launch_step_tb<TB, 256>(g, s); break;
    }
}

void launch_step(const Group* g, hipStream_t s) {
    if (g->model == DBA_MODEL_DENDRITE_K94 || g->model == DBA_MODEL_MODELB) {
        dim3 grid(g->grid), block(64);
        if (g->model == DBA_MODEL_MODELB) {
            // each pair returns at once when the other takes the step (by list length);
            // superblocks of 2 x 2 x 1 blocks (2 x 2 x 2 on 8-wave workgroups measured slower)
            const int from = g->sb_from > 0 ? g->sb_from : INT_MAX;
            const dim3 sgrid((unsigned)std::max<int64_t>(1, g->total_tiles / 4)), sblock(Sb<1>::NT);
            if (g->tile_z == 4) {
                if (g->sb_from > 0) hipLaunchKernelGGL(modelb_mu_sb_kernel<1>, sgrid, sblock, 0, s, wl_of(g), from);
                hipLaunchKernelGGL((modelb_mu_kernel<4, 4>), grid, block, 0, s, wl_of(g), from);
                if (g->sb_from > 0)
                    hipLaunchKernelGGL(modelb_update_sb_kernel<1>, sgrid, sblock, 0, s, wl_of(g), from);
                hipLaunchKernelGGL((modelb_update_kernel<4, 4>), grid, block, 0, s, wl_of(g), from);
            } else {
                hipLaunchKernelGGL((modelb_mu_kernel<4, 1>), grid, block, 0, s, wl_of(g), from);
                hipLaunchKernelGGL((modelb_update_kernel<4, 1>), grid, block, 0, s, wl_of(g), from);
            }
        } else {
            if (g->tile_z == 4)
                hipLaunchKernelGGL((k94_step_kernel<4, 4>), grid, block, 0, s, wl_of(g));
            else
                hipLaunchKernelGGL((k94_step_kernel<4, 1>), grid, block, 0, s, wl_of(g));
        }
        return;
    }
    switch (g->tile) {
        case 8: launch_step_nt<8>(g, s); break;
        case 16: launch_step_nt<16>(g, s); break;
        default: launch_step_nt<32>(g, s); break;
    }
}

// Single-launch steps: one instance in one segment (its whole state fits one
// canonicalising workgroup), a model with a single-launch kernel, no conserved
// shift (its per-step global sum needs the finalize).
// Conserved DBA needs the previous step's global change sum: single-launch
// steps recompute it in every workgroup, for one CH instance of <= 1024 tiles.
bool single_launch_params(const dba_params& p, int64_t nt) {
    if (p.mode == DBA_MODE_CONSERVED) return p.model == DBA_MODEL_CH && nt <= kSegTiles;
    return p.model == DBA_MODEL_CH || p.model == DBA_MODEL_DENDRITE_K93;
}

// Groups every instance of which is CH or K93 (one instance in one segment, or a
// tile map for the rest), conserved DBA only as one CH instance in one segment;
// batches only in mode 2, since every sleeping tile still takes a workgroup.
bool single_launch_ok(const dba_ctx* ctx, const Group* g) {
    if (ctx->single_launch == 0 || (ctx->single_launch == 1 && g->ids.size() != 1)) return false;
    if (!(g->ids.size() == 1 && g->nseg == 1) && !g->d_dmap) return false;
    for (int id : g->ids) {
        const Instance* in = ctx->inst.at(id);
        if (!single_launch_params(in->p, in->h.nt)) return false;
        if (in->p.mode == DBA_MODE_CONSERVED && !(g->ids.size() == 1 && g->nseg == 1)) return false;
    }
    return true;
}

template <int K>
void launch_march(const Group* g, hipStream_t s, int first, int step) {
    const dim3 grid((unsigned)(g->total_tiles / K));
    if (g->nine)
        hipLaunchKernelGGL((ch_march_kernel<K, true>), grid, dim3(256), 0, s, g->d_insts, first, step);
    else
        hipLaunchKernelGGL((ch_march_kernel<K, false>), grid, dim3(256), 0, s, g->d_insts, first, step);
}

template <int TB, int NT>
void launch_dense_tb(const Group* g, hipStream_t s, int first, int step) {
    const dim3 grid((unsigned)g->total_tiles);
    if (g->model == DBA_MODEL_CH) {
        constexpr int NTC = NT <= TB * TB ? NT : TB * TB;
        if (g->conserved && g->partial)
            hipLaunchKernelGGL((ch_dense_kernel<TB, NTC, true, true>), grid, dim3(NTC), 0, s, g->d_insts, g->d_dmap,
                               first, step);
        else if (g->conserved)
            hipLaunchKernelGGL((ch_dense_kernel<TB, NTC, true, false>), grid, dim3(NTC), 0, s, g->d_insts, g->d_dmap,
                               first, step);
        else if (g->partial)
            hipLaunchKernelGGL((ch_dense_kernel<TB, NTC, false, true>), grid, dim3(NTC), 0, s, g->d_insts, g->d_dmap,
                               first, step);
        else
            hipLaunchKernelGGL((ch_dense_kernel<TB, NTC, false, false>), grid, dim3(NTC), 0, s, g->d_insts,
                               g->d_dmap, first, step);
    } else if (g->partial) {
        hipLaunchKernelGGL((k93_dense_kernel<TB, NT, true>), grid, dim3(NT), 0, s, g->d_insts, g->d_dmap, first,
                           step);
    } else if (TB == 16 && NT == 512 && g->spec) {
        hipLaunchKernelGGL((k93_dense_kernel<16, 512, false, true>), grid, dim3(512), 0, s, g->d_insts, g->d_dmap,
                           first, step);
    } else {
        hipLaunchKernelGGL((k93_dense_kernel<TB, NT, false>), grid, dim3(NT), 0, s, g->d_insts, g->d_dmap, first,
                           step);
    }
}

template <int TB>
void launch_dense_nt(const Group* g, hipStream_t s, int first, int step) {
    switch (g->dense_threads) {
        case 64: launch_dense_tb<TB, 64>(g, s, first, step); break;
        case 128: launch_dense_tb<TB, 128>(g, s, first, step); break;
        case 512: launch_dense_tb<TB, 512>(g, s, first, step); break;
        case 1024: launch_dense_tb<TB, 1024>(g, s, first, step); break;
        default: launch_dense_tb<TB, 256>(g, s, first, step); break;
    }
}

void launch_dense(const Group* g, hipStream_t s, int first, int step) {
    if (g->march) {
        switch (g->march) {
            case 32: launch_march<32>(g, s, first, step); break;
            case 16: launch_march<16>(g, s, first, step); break;
            case 8: launch_march<8>(g, s, first, step); break;
            case 4: launch_march<4>(g, s, first, step); break;
            default: launch_march<2>(g, s, first, step); break;
        }
        return;
    }
    switch (g->tile) {
        case 8: launch_dense_nt<8>(g, s, first, step); break;
        case 16: launch_dense_nt<16>(g, s, first, step); break;
        default: launch_dense_nt<32>(g, s, first, step); break;
    }
}

void launch_finalize(const Group* g, hipStream_t s, int rotate) {
    hipLaunchKernelGGL(finalize_kernel, dim3((unsigned)g->nseg), dim3(kFinNT), 0, s, g->d_insts, g->d_segs,
                       g->nseg, g->d_parts, g->d_counts, g->d_offsets, rotate, g->ids.size() == 1 ? 1 : 0);
    if (g->nseg > 1) {
        // workgroup 0 scans the segment counts; the others reduce 16 instances each
        const int n_inst = (int)g->ids.size();
        const unsigned wgs = 1u + (unsigned)((n_inst + kFinNT / 64 - 1) / (kFinNT / 64));
        hipLaunchKernelGGL(finalize_merge_kernel, dim3(wgs), dim3(kFinNT), 0, s, g->d_insts, g->nseg, g->d_isegs,
                           n_inst, g->d_parts, g->d_counts, g->d_offsets, rotate);
    }
}

hipEvent_t pooled_event(dba_ctx* ctx) {
    hipEvent_t e = nullptr;
    if (!ctx->ev_pool.empty()) {
        e = ctx->ev_pool.back();
        ctx->ev_pool.pop_back();
    } else {
        (void)hipEventCreate(&e);
    }
    ctx->ev_used.push_back(e);
    return e;
}

dba_status partition(dba_ctx* ctx, const int32_t* ids, int32_t n_ids, std::vector<std::vector<int>>* parts) {
    if (n_ids <= 0 || !ids) return fail(DBA_ERR_INVALID, "no instances given");
    std::vector<int> seen;
    for (int k = 0; k < n_ids; k++) {
        int id = ids[k];
        auto it = ctx->inst.find(id);
        if (it == ctx->inst.end()) return fail(DBA_ERR_NOTFOUND, "unknown instance %d", id);
        if (std::find(seen.begin(), seen.end(), id) != seen.end())
            return fail(DBA_ERR_INVALID, "instance %d listed twice", id);
        seen.push_back(id);
        const dba_params& p = it->second->p;
        int thr = p.threads ? p.threads : default_threads(p.model, p.tile_x);
        bool placed = false;
        for (auto& part : *parts) {
            const dba_params& q = ctx->inst[part[0]]->p;
            int thq = q.threads ? q.threads : default_threads(q.model, q.tile_x);
            if (q.model == p.model && q.tile_x == p.tile_x && q.tile_z == p.tile_z && thq == thr &&
                part.size() < 255) {
                part.push_back(id);
                placed = true;
                break;
            }
        }
        if (!placed) parts->push_back({id});
    }
    return DBA_OK;
}

dba_status enqueue_steps(dba_ctx* ctx, const std::vector<Group*>& gs, int n_steps, bool timed) {
    // single-launch groups: n steps, then one canonicalisation
    for (Group* g : gs) {
        if (n_steps == 0 || !single_launch_ok(ctx, g)) continue;
        for (int s = 0; s < n_steps + 1; s++) {
            hipEvent_t a = nullptr, b = nullptr;
            if (timed) {
                a = pooled_event(ctx);
                b = pooled_event(ctx);
                (void)hipEventRecord(a, ctx->stream);
            }
            if (s < n_steps) {
                launch_dense(g, ctx->stream, s == 0, s);
            } else if (!g->d_dmap) {
                hipLaunchKernelGGL(dense_canon_kernel, dim3(1), dim3(kFinNT), 0, ctx->stream, g->d_insts, n_steps);
            } else {
                hipLaunchKernelGGL(dense_canon_seg_kernel, dim3((unsigned)g->nseg), dim3(kFinNT), 0, ctx->stream,
                                   g->d_insts, g->d_segs, g->d_parts, g->d_counts, n_steps);
                LAUNCH_CHECK();
                hipLaunchKernelGGL(dense_canon_merge_kernel, dim3(1), dim3(kFinNT), 0, ctx->stream, g->d_insts,
                                   g->d_isegs, (int)g->ids.size(), g->d_parts, g->d_counts, n_steps);
            }
            LAUNCH_CHECK();
            if (timed) {
                (void)hipEventRecord(b, ctx->stream);
                (s < n_steps ? ctx->ev_step : ctx->ev_fin).push_back({a, b});
            }
        }
    }
    for (int s = 0; s < n_steps; s++) {
        for (Group* g : gs) {
            if (single_launch_ok(ctx, g)) continue;
            hipEvent_t a = nullptr, b = nullptr, c = nullptr;
            if (timed) {
                a = pooled_event(ctx);
                b = pooled_event(ctx);
                c = pooled_event(ctx);
                (void)hipEventRecord(a, ctx->stream);
            }
            launch_step(g, ctx->stream);
            LAUNCH_CHECK();
            if (timed) (void)hipEventRecord(b, ctx->stream);
            launch_finalize(g, ctx->stream, 1);
            LAUNCH_CHECK();
            if (timed) {
                (void)hipEventRecord(c, ctx->stream);
                ctx->ev_step.push_back({a, b});
                ctx->ev_fin.push_back({b, c});
            }
        }
    }
    return DBA_OK;
}

dba_status prepare_groups(dba_ctx* ctx, const int32_t* ids, int32_t n_ids, std::vector<Group*>* gs) {
    std::vector<std::vector<int>> parts;
    dba_status st = partition(ctx, ids, n_ids, &parts);
    if (st != DBA_OK) return st;
    ctx->clock++;
    for (int k = 0; k < n_ids; k++) ctx->inst[ids[k]]->stepped = true;
    for (auto& part : parts) {
        Group* g = nullptr;
        st = get_group(ctx, part, &g);
        if (st != DBA_OK) return st;
        gs->push_back(g);
    }
    // KH instances whose state was uploaded since their last step
    for (int k = 0; k < n_ids; k++) {
        Instance* in = ctx->inst[ids[k]];
        if (in->needs_prime) {
            st = kh_prime(ctx, in);
            if (st != DBA_OK) return st;
        }
    }
    // (re)build the worklists from the current awake sets
    for (Group* g : *gs) {
        launch_finalize(g, ctx->stream, 0);
        LAUNCH_CHECK();
    }
    return DBA_OK;
}

Instance* find(dba_ctx* ctx, int id) {
    auto it = ctx->inst.find(id);
    return it == ctx->inst.end() ? nullptr : it->second;
}

dba_status ensure_staging(dba_ctx* ctx, size_t n) {
    if (ctx->staging_n >= n) return DBA_OK;
    if (ctx->staging) (void)hipFree(ctx->staging);
    ctx->staging = nullptr;
    ctx->staging_n = 0;
    HIP_TRY(hipMalloc(&ctx->staging, n * sizeof(double)));
    ctx->staging_n = n;
    return DBA_OK;
}

}  // namespace

// ===================================================================== C ABI

namespace {
// Device scratch of one droplet analysis, freed when the call returns.
struct CcScratch {
    std::vector<void*> ptrs;
    ~CcScratch() {
        for (void* q : ptrs) (void)hipFree(q);
    }
    template <typename T>
    hipError_t get(T** out, size_t count) {
        void* q = nullptr;
        const hipError_t e = hipMalloc(&q, sizeof(T) * (count ? count : 1));
        if (e == hipSuccess) ptrs.push_back(q);
        *out = (T*)q;
        return e;
    }
};
}  // namespace

extern "C" {

const char* dba_last_error(void) { return g_error.c_str(); }

dba_status dba_init(int32_t device, dba_ctx** out) {
    if (!out) return fail(DBA_ERR_INVALID, "out is null");
    int n = 0;
    HIP_TRY(hipGetDeviceCount(&n));
    if (device < 0 || device >= n) return fail(DBA_ERR_INVALID, "device %d not present (%d devices)", device, n);
    HIP_TRY(hipSetDevice(device));
    dba_ctx* ctx = new dba_ctx();
    ctx->device = device;
    hipError_t e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete ctx;
        return fail(DBA_ERR_HIP, "hipStreamCreate failed: %s", hipGetErrorString(e));
    }
    e = hipStreamCreateWithFlags(&ctx->setup, hipStreamNonBlocking);
    if (e != hipSuccess) {
        (void)hipStreamDestroy(ctx->stream);
        delete ctx;
        return fail(DBA_ERR_HIP, "hipStreamCreate failed: %s", hipGetErrorString(e));
    }
    void* sp = nullptr;
    e = hipGetSymbolAddress(&sp, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !sp) {
        (void)hipStreamDestroy(ctx->stream);
        (void)hipStreamDestroy(ctx->setup);
        delete ctx;
        return fail(DBA_ERR_HIP, "hipGetSymbolAddress(scratch) failed: %s", hipGetErrorString(e));
    }
    ctx->scratch = (unsigned char*)sp;
    HIP_TRY(hipMalloc(&ctx->d_flag, sizeof(int)));
    HIP_TRY(hipMalloc(&ctx->d_diag, sizeof(double) * (3 * kDiagBlocks + kMaxFrontCols + 8)));
    HIP_TRY(hipMalloc(&ctx->d_cols, sizeof(int32_t) * kMaxFrontCols));
    for (auto& m : ctx->markers) HIP_TRY(hipEventCreate(&m));
    *out = ctx;
    return DBA_OK;
}

dba_status dba_finalize(dba_ctx* ctx) {
    if (!ctx) return DBA_OK;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    std::vector<int> ids;
    for (auto& kv : ctx->inst) ids.push_back(kv.first);
    for (int id : ids) dba_instance_destroy(ctx, id);
    for (auto& kv : ctx->graphs) (void)hipGraphExecDestroy(kv.second.exec);
    for (auto e : ctx->ev_used) (void)hipEventDestroy(e);
    for (auto e : ctx->ev_pool) (void)hipEventDestroy(e);
    for (auto m : ctx->markers) (void)hipEventDestroy(m);
    if (ctx->staging) (void)hipFree(ctx->staging);
    if (ctx->d_flag) (void)hipFree(ctx->d_flag);
    if (ctx->d_diag) (void)hipFree(ctx->d_diag);
    if (ctx->d_cols) (void)hipFree(ctx->d_cols);
    (void)hipStreamDestroy(ctx->stream);
    (void)hipStreamDestroy(ctx->setup);
    delete ctx;
    return DBA_OK;
}

dba_status dba_footprint(const dba_params* p, int64_t* bytes) {
    if (!p || !bytes) return fail(DBA_ERR_INVALID, "null argument");
    dba_status st = validate(*p);
    if (st != DBA_OK) return st;
    *bytes = footprint_of(*p);
    return DBA_OK;
}

dba_status dba_device_memory(dba_ctx* ctx, int64_t* free_bytes, int64_t* total_bytes) {
    if (!ctx || !free_bytes || !total_bytes) return fail(DBA_ERR_INVALID, "null argument");
    HIP_TRY(hipSetDevice(ctx->device));
    size_t f = 0, t = 0;
    HIP_TRY(hipMemGetInfo(&f, &t));
    *free_bytes = (int64_t)f;
    *total_bytes = (int64_t)t;
    return DBA_OK;
}

dba_status dba_instance_create(dba_ctx* ctx, const dba_params* pp, int32_t* id) {
    if (!ctx || !pp || !id) return fail(DBA_ERR_INVALID, "null argument");
    dba_status st = validate(*pp);
    if (st != DBA_OK) return st;
    HIP_TRY(hipSetDevice(ctx->device));
    Instance* in = new Instance();
    in->p = *pp;
    if (in->p.tile_z <= 0) in->p.tile_z = 1;
    const dba_params& p = in->p;
    Layout L = layout_of(p);
    Inst& h = in->h;
    memset(&h, 0, sizeof(h));
    h.model = p.model;
    h.mode = p.mode;
    h.nx = p.nx; h.ny = p.ny; h.nz = p.nz;
    h.tx = p.tile_x; h.ty = p.tile_y; h.tz = p.tile_z;
    h.ntx = L.ntx; h.nty = L.nty; h.ntz = L.ntz; h.nt = L.nt; h.nwords = L.nwords;
    h.bc_x = p.bc_x; h.bc_y = p.bc_y; h.bc_z = p.bc_z;
    h.stencil = p.stencil; h.aniso_j = p.aniso_j; h.dt_mode = p.dt_mode; h.metric_kind = p.metric_kind;
    h.nsets = L.nsets;
    h.nfields = L.nf;
    h.ncons = 0;
    if (p.mode == DBA_MODE_CONSERVED) {
        if (p.model == DBA_MODEL_CH) h.ncons = 1;
        if (p.model == DBA_MODEL_KH) h.ncons = 4;
        if (p.model == DBA_MODEL_MODELB) h.ncons = 3;
    }
    for (int f = 0; f < kMaxFields; f++) h.field_set[f] = 0;
    if (L.nsets == 2) h.field_set[1] = 1;
    h.full_tiles = (p.nx % p.tile_x == 0 && p.ny % p.tile_y == 0 && p.nz % p.tile_z == 0) ? 1 : 0;
    h.dx = p.dx; h.dy = p.dy > 0 ? p.dy : p.dx; h.dz = p.dz > 0 ? p.dz : p.dx; h.dt = p.dt;
    h.thresh_a = p.thresh_a; h.thresh_b = p.thresh_b;
    h.ntotal = (double)L.n;
    derive_constants(p, h);
    h.tau = p.tau; h.eps_bar = p.eps_bar; h.delta = p.delta; h.alpha = p.alpha; h.gamma = p.gamma;
    h.theta0 = p.theta0; h.latent_k = p.latent_k; h.t_eq = p.t_eq;
    h.mobility = p.mobility; h.kappa = p.kappa; h.react_gamma = p.react_gamma; h.cbar = p.cbar;
    h.chi_xy = p.chi_xy; h.chi_xz = p.chi_xz; h.chi_yz = p.chi_yz;
    h.chi_xb = p.chi_xb; h.chi_yb = p.chi_yb; h.chi_zb = p.chi_zb;
    h.v_x = p.v_x; h.v_y = p.v_y; h.v_z = p.v_z; h.r_x = p.r_x; h.r_y = p.r_y; h.r_z = p.r_z;
    h.n_exp = p.n_exp; h.m_exp = p.m_exp;
    h.epsx_sq = p.epsx_sq; h.epsy_sq = p.epsy_sq; h.epsz_sq = p.epsz_sq; h.epsphi_sq = p.epsphi_sq;
    h.mob_x = p.mob_x; h.mob_y = p.mob_y; h.mob_z = p.mob_z; h.mob_phi = p.mob_phi;
    h.k_eq = p.k_eq; h.z_crit = p.z_crit; h.p_gel = p.p_gel; h.k_0 = p.k_0;
    h.gas_gamma = p.gas_gamma; h.cfl = p.cfl; h.vol = h.dx * h.dy;
    h.hist_cap = std::max(0, p.history_cap);

#define ALLOC(ptr, count)                         \
    do {                                          \
        dba_status s_ = dalloc(in, &(ptr), (count), ctx->setup); \
        if (s_ != DBA_OK) { free_instance(in); delete in; return s_; } \
    } while (0)
    for (int f = 0; f < L.nf; f++) {
        ALLOC(h.buf[0][f], (size_t)L.n);
        ALLOC(h.buf[1][f], (size_t)L.n);
    }
    for (int a = 0; a < L.naux; a++) ALLOC(h.aux[a], (size_t)L.aux_len);
    for (int s = 0; s < 2; s++) {
        ALLOC(h.a_cur[s], (size_t)L.nwords);
        ALLOC(h.nflag[s], (size_t)L.nt);
        ALLOC(h.par[s], (size_t)L.nwords);
        ALLOC(h.dF[s], (size_t)L.nsets * L.nt);
        ALLOC(h.dP[s], (size_t)L.nsets * L.nt);
    }
    ALLOC(h.druns, (size_t)L.nt);
    ALLOC(h.tpart, (size_t)kMaxFields * L.nt);
    ALLOC(h.tspeed, (size_t)2 * L.nt);
    ALLOC(h.s_tile, (size_t)kMaxFields * 2 * L.nt);
    ALLOC(h.scal, 16);
    ALLOC(h.ctr, 4);
    ALLOC(h.hist, (size_t)2 * h.hist_cap + 2);
    ALLOC(h.seg, (size_t)L.nt);
    ALLOC(h.nbrec, (size_t)L.nt);
    ALLOC(in->d_self, 1);
#undef ALLOC
    // all tiles awake at start (example_Dendritic_growth/kernel.cu:149-161)
    std::vector<uint64_t> all((size_t)L.nwords, ~0ull);
    if (L.nt % 64) all.back() = (1ull << (L.nt % 64)) - 1ull;
    double scal[16] = {0};
    scal[1] = p.dt;
    for (int s = 0; s < L.nsets; s++)
        HIP_TRY(hipMemcpyAsync(h.a_cur[s], all.data(), sizeof(uint64_t) * L.nwords, hipMemcpyHostToDevice,
                               ctx->setup));
    HIP_TRY(hipMemcpyAsync(h.scal, scal, sizeof(scal), hipMemcpyHostToDevice, ctx->setup));
    HIP_TRY(hipMemcpyAsync(in->d_self, &h, sizeof(Inst), hipMemcpyHostToDevice, ctx->setup));
    // set-up lands before any later work on the compute stream (host-side wait);
    // work already running there for other instances is not waited for
    HIP_TRY(hipStreamSynchronize(ctx->setup));
    in->footprint = footprint_of(p);
    in->ncell = L.n;
    in->id = ctx->next_id++;
    ctx->inst[in->id] = in;
    *id = in->id;
    return DBA_OK;
}

dba_status dba_instance_destroy(dba_ctx* ctx, int32_t id) {
    if (!ctx) return fail(DBA_ERR_INVALID, "null ctx");
    Instance* in = find(ctx, id);
    if (!in) return fail(DBA_ERR_NOTFOUND, "unknown instance %d", id);
    (void)hipStreamSynchronize(ctx->stream);
    invalidate_groups(ctx, id);
    free_instance(in);
    ctx->inst.erase(id);
    delete in;
    return DBA_OK;
}

dba_status dba_instance_info(dba_ctx* ctx, int32_t id, dba_info* out) {
    if (!ctx || !out) return fail(DBA_ERR_INVALID, "null argument");
    Instance* in = find(ctx, id);
    if (!in) return fail(DBA_ERR_NOTFOUND, "unknown instance %d", id);
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    double scal[16];
    int64_t ctr[4];
    HIP_TRY(hipMemcpy(scal, in->h.scal, sizeof(scal), hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(ctr, in->h.ctr, sizeof(ctr), hipMemcpyDeviceToHost));
    out->steps_done = ctr[0];
    out->field_updates = ctr[1];
    out->cell_updates = ctr[1] / std::max(1, in->h.nsets);
    out->ntiles = in->h.nt;
    out->ntiles_x = in->h.ntx;
    out->ntiles_y = in->h.nty;
    out->ntiles_z = in->h.ntz;
    out->sim_time = scal[2];
    out->last_dt = scal[3];
    out->cons_offset = scal[8];
    out->footprint_bytes = in->footprint;
    out->single_launch = ctx->single_launch != 0 && single_launch_params(in->p, in->h.nt) ? 1 : 0;
    out->reserved = 0;
    return DBA_OK;
}

dba_status dba_upload(dba_ctx* ctx, int32_t id, int32_t field, const double* host, size_t n) {
    if (!ctx || !host) return fail(DBA_ERR_INVALID, "null argument");
    Instance* in = find(ctx, id);
    if (!in) return fail(DBA_ERR_NOTFOUND, "unknown instance %d", id);
    if (field < 0 || field >= in->h.nfields) return fail(DBA_ERR_INVALID, "field %d invalid", field);
    if ((int64_t)n != in->ncell) return fail(DBA_ERR_INVALID, "n=%zu, expected %lld", n, (long long)in->ncell);
    // An instance that was never stepped is touched by no queued kernel, so its
    // upload overlaps the steps of other instances; otherwise it is ordered
    // after them on the compute stream.
    hipStream_t s = in->stepped ? ctx->stream : ctx->setup;
    HIP_TRY(hipMemcpyAsync(in->h.buf[0][field], host, n * sizeof(double), hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(in->h.buf[1][field], in->h.buf[0][field], n * sizeof(double), hipMemcpyDeviceToDevice, s));
    if (in->h.ncons > 0) {
        hipLaunchKernelGGL(reset_epochs_kernel, dim3(64), dim3(256), 0, s, in->d_self, 0);
        LAUNCH_CHECK();
    }
    if (in->p.model == DBA_MODEL_KH) in->needs_prime = true;
    HIP_TRY(hipStreamSynchronize(s));
    return DBA_OK;
}

dba_status dba_download(dba_ctx* ctx, int32_t id, int32_t field, double* host, size_t n) {
    if (!ctx || !host) return fail(DBA_ERR_INVALID, "null argument");
    Instance* in = find(ctx, id);
    if (!in) return fail(DBA_ERR_NOTFOUND, "unknown instance %d", id);
    if (field < 0 || field >= in->h.nfields) return fail(DBA_ERR_INVALID, "field %d invalid", field);
    if ((int64_t)n != in->ncell) return fail(DBA_ERR_INVALID, "n=%zu, expected %lld", n, (long long)in->ncell);
    dba_status st = ensure_staging(ctx, n);
    if (st != DBA_OK) return st;
    unsigned blocks = (unsigned)std::min<int64_t>(4096, (in->ncell + 255) / 256);
    hipLaunchKernelGGL(materialize_kernel, dim3(blocks), dim3(256), 0, ctx->stream, in->d_self, 0, field,
                       ctx->staging);
    LAUNCH_CHECK();
    HIP_TRY(hipMemcpyAsync(host, ctx->staging, n * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return DBA_OK;
}

dba_status dba_step(dba_ctx* ctx, const int32_t* ids, int32_t n_ids, int32_t n_steps) {
    if (!ctx) return fail(DBA_ERR_INVALID, "null ctx");
    if (n_steps < 0) return fail(DBA_ERR_INVALID, "n_steps < 0");
    std::vector<Group*> gs;
    dba_status st = prepare_groups(ctx, ids, n_ids, &gs);
    if (st != DBA_OK) return st;
    return enqueue_steps(ctx, gs, n_steps, ctx->timing);
}

dba_status dba_step_graph(dba_ctx* ctx, const int32_t* ids, int32_t n_ids, int32_t n_steps, int32_t chunk) {
    if (!ctx) return fail(DBA_ERR_INVALID, "null ctx");
    if (n_steps < 0 || chunk < 1) return fail(DBA_ERR_INVALID, "n_steps < 0 or chunk < 1");
    std::vector<Group*> gs;
    dba_status st = prepare_groups(ctx, ids, n_ids, &gs);
    if (st != DBA_OK) return st;
    int reps = n_steps / chunk, rem = n_steps % chunk;
    if (reps > 0) {
        GraphKey key{std::vector<int>(ids, ids + n_ids), chunk, ctx->single_launch};
        auto it = ctx->graphs.find(key);
        hipGraphExec_t exec = nullptr;
        if (it == ctx->graphs.end()) {
            evict_graphs(ctx);
            hipGraph_t graph = nullptr;
            HIP_TRY(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeRelaxed));
            dba_status est = enqueue_steps(ctx, gs, chunk, false);
            hipError_t e = hipStreamEndCapture(ctx->stream, &graph);
            if (est != DBA_OK) return est;
            if (e != hipSuccess) return fail(DBA_ERR_HIP, "hipStreamEndCapture: %s", hipGetErrorString(e));
            HIP_TRY(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
            (void)hipGraphDestroy(graph);
            ctx->graphs[key] = CachedGraph{exec, ctx->clock};
        } else {
            it->second.last_use = ctx->clock;
            exec = it->second.exec;
        }
        for (int r = 0; r < reps; r++) HIP_TRY(hipGraphLaunch(exec, ctx->stream));
    }
    if (rem > 0) return enqueue_steps(ctx, gs, rem, false);
    return DBA_OK;
}

dba_status dba_set_single_launch(dba_ctx* ctx, int32_t mode) {
    if (!ctx) return fail(DBA_ERR_INVALID, "null ctx");
    if (mode < 0 || mode > 2) return fail(DBA_ERR_INVALID, "single-launch mode must be 0, 1 or 2 (got %d)", mode);
    // captured graphs are cached per mode, so switching keeps them
    ctx->single_launch = mode;
    return DBA_OK;
}

dba_status dba_sync(dba_ctx* ctx) {
    if (!ctx) return fail(DBA_ERR_INVALID, "null ctx");
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return DBA_OK;
}

dba_status dba_active_tiles(dba_ctx* ctx, int32_t id, int32_t set, int32_t* count, uint8_t* mask) {
    if (!ctx || !count) return fail(DBA_ERR_INVALID, "null argument");
    Instance* in = find(ctx, id);
    if (!in) return fail(DBA_ERR_NOTFOUND, "unknown instance %d", id);
    if (set < 0 || set >= in->h.nsets) return fail(DBA_ERR_INVALID, "set %d invalid", set);
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    int nt = in->h.nt;
    std::vector<uint64_t> w((size_t)in->h.nwords);
    HIP_TRY(hipMemcpy(w.data(), in->h.a_cur[set], sizeof(uint64_t) * w.size(), hipMemcpyDeviceToHost));
    int c = 0;
    for (int t = 0; t < nt; t++) {
        int b = in->p.mode == DBA_MODE_REGULAR ? 1 : (int)((w[t >> 6] >> (t & 63)) & 1ull);
        c += b;
        if (mask) mask[t] = (uint8_t)b;
    }
    *count = c;
    return DBA_OK;
}

dba_status dba_worklist(dba_ctx* ctx, int32_t id, int32_t* tiles, int32_t cap, int32_t* count) {
    if (!ctx || !count) return fail(DBA_ERR_INVALID, "null argument");
    Instance* in = find(ctx, id);
    if (!in) return fail(DBA_ERR_NOTFOUND, "unknown instance %d", id);
    // compact the instance's current awake set (the list the next step runs)
    std::vector<Group*> gs;
    dba_status st = prepare_groups(ctx, &id, 1, &gs);
    if (st != DBA_OK) return st;
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    int64_t ctr[4];
    HIP_TRY(hipMemcpy(ctr, in->h.ctr, sizeof(ctr), hipMemcpyDeviceToHost));
    *count = (int32_t)ctr[2];
    if (tiles && cap > 0) {
        int n = (int)std::min<int64_t>(cap, ctr[2]);
        if (n > 0) HIP_TRY(hipMemcpy(tiles, in->h.seg, sizeof(int32_t) * n, hipMemcpyDeviceToHost));
    }
    return DBA_OK;
}

dba_status dba_active_history(dba_ctx* ctx, int32_t id, int32_t* counts, int32_t max_steps, int32_t* n_steps) {
    if (!ctx || !n_steps) return fail(DBA_ERR_INVALID, "null argument");
    Instance* in = find(ctx, id);
    if (!in) return fail(DBA_ERR_NOTFOUND, "unknown instance %d", id);
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    int64_t ctr[4];
    HIP_TRY(hipMemcpy(ctr, in->h.ctr, sizeof(ctr), hipMemcpyDeviceToHost));
    int n = (int)std::min<int64_t>(std::min<int64_t>(ctr[0], in->h.hist_cap), max_steps);
    if (counts && n > 0) HIP_TRY(hipMemcpy(counts, in->h.hist, sizeof(int32_t) * 2 * n, hipMemcpyDeviceToHost));
    *n_steps = n;
    return DBA_OK;
}

dba_status dba_timing_enable(dba_ctx* ctx, int32_t enable) {
    if (!ctx) return fail(DBA_ERR_INVALID, "null ctx");
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    ctx->ev_pool.insert(ctx->ev_pool.end(), ctx->ev_used.begin(), ctx->ev_used.end());
    ctx->ev_used.clear();
    ctx->ev_step.clear();
    ctx->ev_fin.clear();
    ctx->timing = enable != 0;
    return DBA_OK;
}

dba_status dba_timing_query(dba_ctx* ctx, double* step_ms, int64_t* step_n, double* fin_ms, int64_t* fin_n) {
    if (!ctx) return fail(DBA_ERR_INVALID, "null ctx");
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    double a = 0.0, b = 0.0;
    for (auto& e : ctx->ev_step) {
        float ms = 0.f;
        HIP_TRY(hipEventElapsedTime(&ms, e.first, e.second));
        a += ms;
    }
    for (auto& e : ctx->ev_fin) {
        float ms = 0.f;
        HIP_TRY(hipEventElapsedTime(&ms, e.first, e.second));
        b += ms;
    }
    if (step_ms) *step_ms = a;
    if (step_n) *step_n = (int64_t)ctx->ev_step.size();
    if (fin_ms) *fin_ms = b;
    if (fin_n) *fin_n = (int64_t)ctx->ev_fin.size();
    return DBA_OK;
}

dba_status dba_marker(dba_ctx* ctx, int32_t slot) {
    if (!ctx || slot < 0 || slot >= 8) return fail(DBA_ERR_INVALID, "bad marker slot");
    HIP_TRY(hipEventRecord(ctx->markers[slot], ctx->stream));
    return DBA_OK;
}

dba_status dba_marker_elapsed(dba_ctx* ctx, int32_t a, int32_t b, double* ms) {
    if (!ctx || !ms || a < 0 || a >= 8 || b < 0 || b >= 8) return fail(DBA_ERR_INVALID, "bad marker slot");
    HIP_TRY(hipEventSynchronize(ctx->markers[b]));
    float f = 0.f;
    HIP_TRY(hipEventElapsedTime(&f, ctx->markers[a], ctx->markers[b]));
    *ms = f;
    return DBA_OK;
}

dba_status dba_check_finite(dba_ctx* ctx, int32_t id) {
    if (!ctx) return fail(DBA_ERR_INVALID, "null ctx");
    Instance* in = find(ctx, id);
    if (!in) return fail(DBA_ERR_NOTFOUND, "unknown instance %d", id);
    HIP_TRY(hipMemsetAsync(ctx->d_flag, 0, sizeof(int), ctx->stream));
    unsigned blocks = (unsigned)std::min<int64_t>(2048, (in->ncell + 255) / 256);
    hipLaunchKernelGGL(finite_kernel, dim3(blocks), dim3(256), 0, ctx->stream, in->d_self, 0, ctx->d_flag);
    LAUNCH_CHECK();
    int bad = 0;
    HIP_TRY(hipMemcpyAsync(&bad, ctx->d_flag, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    if (bad) return fail(DBA_ERR_NUMERIC, "instance %d holds non-finite values", id);
    return DBA_OK;
}

dba_status dba_field_sum(dba_ctx* ctx, int32_t id, int32_t field, double* sum) {
    if (!ctx || !sum) return fail(DBA_ERR_INVALID, "null argument");
    Instance* in = find(ctx, id);
    if (!in) return fail(DBA_ERR_NOTFOUND, "unknown instance %d", id);
    if (field < 0 || field >= in->h.nfields) return fail(DBA_ERR_INVALID, "field %d invalid", field);
    double* part = ctx->d_diag;
    double* res = ctx->d_diag + 3 * kDiagBlocks;
    hipLaunchKernelGGL(field_sum_kernel, dim3(kDiagBlocks), dim3(kDiagNT), 0, ctx->stream, in->d_self, 0, field,
                       part);
    LAUNCH_CHECK();
    hipLaunchKernelGGL(reduce_parts_kernel, dim3(1), dim3(64), 0, ctx->stream, part, 1, res);
    LAUNCH_CHECK();
    HIP_TRY(hipMemcpyAsync(sum, res, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return DBA_OK;
}

dba_status dba_growth_front(dba_ctx* ctx, int32_t id, const int32_t* xs, int32_t n, double* front) {
    if (!ctx || !xs || !front) return fail(DBA_ERR_INVALID, "null argument");
    Instance* in = find(ctx, id);
    if (!in) return fail(DBA_ERR_NOTFOUND, "unknown instance %d", id);
    if (in->p.model != DBA_MODEL_DENDRITE_K93 && in->p.model != DBA_MODEL_DENDRITE_K94)
        return fail(DBA_ERR_INVALID, "growth front needs a dendrite instance");
    if (n < 0 || n > kMaxFrontCols) return fail(DBA_ERR_INVALID, "n must be in [0, %d]", kMaxFrontCols);
    for (int k = 0; k < n; k++)
        if (xs[k] < 0 || xs[k] >= in->p.nx) return fail(DBA_ERR_INVALID, "column %d outside the grid", xs[k]);
    if (n == 0) return DBA_OK;
    // synchronous so the host array may be reused at once; ordered after queued steps
    HIP_TRY(hipMemcpyAsync(ctx->d_cols, xs, sizeof(int32_t) * n, hipMemcpyHostToDevice, ctx->stream));
    double* res = ctx->d_diag + 3 * kDiagBlocks;
    hipLaunchKernelGGL(growth_front_kernel, dim3(1), dim3(kMaxFrontCols), 0, ctx->stream, in->d_self, 0,
                       ctx->d_cols, n, res);
    LAUNCH_CHECK();
    HIP_TRY(hipMemcpyAsync(front, res, sizeof(double) * n, hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return DBA_OK;
}

dba_status dba_droplets(dba_ctx* ctx, int32_t id, int32_t field, double* avg, int64_t* volume,
                        int64_t* n_droplets, int64_t* sizes, int64_t cap) {
    if (!ctx || !avg || !volume || !n_droplets) return fail(DBA_ERR_INVALID, "null argument");
    if (cap < 0 || (cap > 0 && !sizes)) return fail(DBA_ERR_INVALID, "sizes must hold cap >= 0 entries");
    Instance* in = find(ctx, id);
    if (!in) return fail(DBA_ERR_NOTFOUND, "unknown instance %d", id);
    if (field < 0 || field >= in->h.nfields) return fail(DBA_ERR_INVALID, "field %d invalid", field);
    const int64_t n = (int64_t)in->p.nx * in->p.ny * in->p.nz;
    if (n >= INT32_MAX) return fail(DBA_ERR_INVALID, "droplet analysis supports grids below 2^31 cells");
    CcScratch sc;
    uint8_t* cls;
    int32_t *parent, *size, *key;
    double *part, *d_avg;
    int64_t *vpart, *counters;
    if (sc.get(&cls, n) != hipSuccess || sc.get(&parent, n) != hipSuccess || sc.get(&size, n) != hipSuccess ||
        sc.get(&key, n) != hipSuccess || sc.get(&part, 2 * kCcBlocks) != hipSuccess || sc.get(&d_avg, 1) != hipSuccess ||
        sc.get(&vpart, kCcBlocks) != hipSuccess || sc.get(&counters, 2) != hipSuccess)
        return fail(DBA_ERR_NOMEM, "droplet analysis: no device memory for %lld cells", (long long)n);
    hipStream_t st = ctx->stream;
    const dim3 g(kCcBlocks), b(kCcNT);
    hipLaunchKernelGGL(cc_minmax_kernel, g, b, 0, st, in->d_self, 0, field, part);
    hipLaunchKernelGGL(cc_threshold_kernel, dim3(1), dim3(64), 0, st, part, d_avg);
    hipLaunchKernelGGL(cc_init_kernel, g, b, 0, st, in->d_self, 0, field, d_avg, cls, parent, size, key);
    hipLaunchKernelGGL(cc_union_kernel, g, b, 0, st, cls, parent, in->p.nx, in->p.ny, in->p.nz);
    hipLaunchKernelGGL(cc_count_kernel, g, b, 0, st, cls, parent, size, key, n, vpart);
    HIP_TRY(hipMemsetAsync(counters, 0, 2 * sizeof(int64_t), st));
    hipLaunchKernelGGL(cc_collect_kernel, g, b, 0, st, size, key, n, vpart, counters, (int64_t*)nullptr,
                       (int64_t)0);
    LAUNCH_CHECK();
    int64_t cnt[2];
    HIP_TRY(hipMemcpyAsync(cnt, counters, sizeof(cnt), hipMemcpyDeviceToHost, st));
    HIP_TRY(hipMemcpyAsync(avg, d_avg, sizeof(double), hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    *n_droplets = cnt[0];
    *volume = cnt[1];
    if (cap == 0 || cnt[0] == 0) return DBA_OK;
    // second pass with room for every droplet, then the reference's order:
    // droplets by their first seed in index order
    int64_t* pairs;
    if (sc.get(&pairs, 2 * cnt[0]) != hipSuccess) return fail(DBA_ERR_NOMEM, "droplet analysis: no device memory");
    HIP_TRY(hipMemsetAsync(counters, 0, sizeof(int64_t), st));
    hipLaunchKernelGGL(cc_collect_kernel, g, b, 0, st, size, key, n, vpart, counters, pairs, cnt[0]);
    LAUNCH_CHECK();
    std::vector<int64_t> h((size_t)(2 * cnt[0]));
    HIP_TRY(hipMemcpyAsync(h.data(), pairs, sizeof(int64_t) * h.size(), hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    std::vector<std::pair<int64_t, int64_t>> v((size_t)cnt[0]);
    for (int64_t k = 0; k < cnt[0]; k++) v[(size_t)k] = {h[(size_t)(2 * k)], h[(size_t)(2 * k + 1)]};
    std::sort(v.begin(), v.end());
    for (int64_t k = 0; k < cnt[0] && k < cap; k++) sizes[k] = v[(size_t)k].second;
    return DBA_OK;
}

dba_status dba_droplets_noalloc(dba_ctx* ctx, int32_t id, int32_t field, double* avg, int64_t* volume,
                                int64_t* n_droplets, int64_t* sizes, int64_t cap) {
    if (!ctx || !avg || !volume || !n_droplets) return fail(DBA_ERR_INVALID, "null argument");
    if (cap < 0 || (cap > 0 && !sizes)) return fail(DBA_ERR_INVALID, "sizes must hold cap >= 0 entries");
    Instance* in = find(ctx, id);
    if (!in) return fail(DBA_ERR_NOTFOUND, "unknown instance %d", id);
    if (field < 0 || field >= in->h.nfields) return fail(DBA_ERR_INVALID, "field %d invalid", field);
    const int64_t n = (int64_t)in->p.nx * in->p.ny * in->p.nz;
    if (n > kScratchMaxCells)
        return fail(DBA_ERR_INVALID,
                    "dba_droplets_noalloc: %lld cells exceed the static scratch of %lld cells; use dba_droplets",
                    (long long)n, (long long)kScratchMaxCells);
    // carve the scratch: every region 256-byte aligned, at most kScratchBytes in all
    unsigned char* q = ctx->scratch;
    auto take = [&q](size_t bytes) {
        unsigned char* r = q;
        q += (bytes + 255) & ~(size_t)255;
        return r;
    };
    uint8_t* cls = take((size_t)n);
    int32_t* parent = (int32_t*)take(4 * (size_t)n);
    int32_t* size = (int32_t*)take(4 * (size_t)n);
    int32_t* key = (int32_t*)take(4 * (size_t)n);
    double* part = (double*)take(sizeof(double) * 2 * kCcBlocks);
    double* d_avg = (double*)take(sizeof(double));
    int64_t* vpart = (int64_t*)take(sizeof(int64_t) * kCcBlocks);
    int64_t* counters = (int64_t*)take(sizeof(int64_t) * 2);
    int32_t* bcount = (int32_t*)take(sizeof(int32_t) * kCcBlocks);
    if ((size_t)(q - ctx->scratch) > kScratchBytes)
        return fail(DBA_ERR_INVALID, "dba_droplets_noalloc: scratch layout exceeds its array");
    hipStream_t st = ctx->stream;
    const dim3 g(kCcBlocks), b(kCcNT);
    hipLaunchKernelGGL(cc_minmax_kernel, g, b, 0, st, in->d_self, 0, field, part);
    hipLaunchKernelGGL(cc_threshold_kernel, dim3(1), dim3(64), 0, st, part, d_avg);
    hipLaunchKernelGGL(cc_init_kernel, g, b, 0, st, in->d_self, 0, field, d_avg, cls, parent, size, key);
    hipLaunchKernelGGL(cc_union_kernel, g, b, 0, st, cls, parent, in->p.nx, in->p.ny, in->p.nz);
    hipLaunchKernelGGL(cc_count_kernel, g, b, 0, st, cls, parent, size, key, n, vpart);
    // the forest is done with: parent[] becomes first[]
    hipLaunchKernelGGL(cc_fill_kernel, g, b, 0, st, parent, n, 0, counters);
    hipLaunchKernelGGL(cc_mark_kernel, g, b, 0, st, size, key, n, vpart, parent, counters);
    LAUNCH_CHECK();
    int64_t cnt[2];
    HIP_TRY(hipMemcpyAsync(cnt, counters, sizeof(cnt), hipMemcpyDeviceToHost, st));
    HIP_TRY(hipMemcpyAsync(avg, d_avg, sizeof(double), hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    *n_droplets = cnt[0];
    *volume = cnt[1];
    if (cap == 0 || cnt[0] == 0) return DBA_OK;
    // the droplets in the order of their first seeds, compacted into size[]
    const int64_t chunk = ((n + kCcBlocks - 1) / kCcBlocks + kCcNT - 1) / kCcNT * kCcNT;
    const int64_t k = std::min<int64_t>(cap, cnt[0]);
    hipLaunchKernelGGL(cc_order_count_kernel, g, b, 0, st, parent, n, chunk, bcount);
    hipLaunchKernelGGL(cc_order_write_kernel, g, b, 0, st, parent, n, chunk, bcount, size, k);
    LAUNCH_CHECK();
    // k int32 sizes into the front of the caller's int64 array, then widened in
    // place from the back (entry j is read before slot j's bytes are written)
    int32_t* narrow = (int32_t*)sizes;
    HIP_TRY(hipMemcpyAsync(narrow, size, sizeof(int32_t) * (size_t)k, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    for (int64_t j = k - 1; j >= 0; j--) sizes[j] = (int64_t)narrow[j];
    return DBA_OK;
}

dba_status dba_kh_vy_mode(dba_ctx* ctx, int32_t id, double* amplitude) {
    if (!ctx || !amplitude) return fail(DBA_ERR_INVALID, "null argument");
    Instance* in = find(ctx, id);
    if (!in) return fail(DBA_ERR_NOTFOUND, "unknown instance %d", id);
    if (in->p.model != DBA_MODEL_KH) return fail(DBA_ERR_INVALID, "vy mode needs a KH instance");
    double* part = ctx->d_diag;
    double* res = ctx->d_diag + 3 * kDiagBlocks;
    hipLaunchKernelGGL(vy_mode_kernel, dim3(kDiagBlocks), dim3(kDiagNT), 0, ctx->stream, in->d_self, 0, part);
    LAUNCH_CHECK();
    hipLaunchKernelGGL(reduce_parts_kernel, dim3(1), dim3(64), 0, ctx->stream, part, 3, res);
    LAUNCH_CHECK();
    double s[3];
    HIP_TRY(hipMemcpyAsync(s, res, sizeof(s), hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    *amplitude = 2.0 * std::sqrt(s[0] * s[0] / (s[2] * s[2]) + s[1] * s[1] / (s[2] * s[2]));
    return DBA_OK;
}

}  // extern "C"
