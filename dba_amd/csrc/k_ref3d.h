// k_ref3d.h - fused DBA steps of the reference's own 3D models on 4x4x4
// activation blocks (4x4x1 when nz == 1), one wave64 per block, periodic.
//
//  * Kobayashi-94 dendrite: Evolve_p(_wakeup) + Evolve_T(_wakeup) + the detector
//    of Wakeup_Next (example_Dendritic_growth/kernel.cu:181-293, 322-574):
//    split p / T sets; regular T = T + dt lapT + p_new - p, DBA T = (T + dt lapT)
//    + (p_new - p) only where p-awake; Pi = 3.14159 (kernel.cu:7).
//  * Ternary Flory-Huggins Model B with reaction: Calc_mu(_wakeup) + detector
//    |lap z| > eps, then Update(_wakeup) + copy-back (example_Reactive_ModelB/
//    kernel.cu:290-763), conserved change capture (example_Reactive_ModelB/
//    Conserved/kernel.cu:630-702). The chemical potentials of sleeping
//    neighbours are the ones they stored when they last ran.
#pragma once
#include "common.h"
#include "mathf64.h"

namespace dba {

struct NbTable3 {
    int tile[27];
    int par[2][27];
    int act[2][27];
    double shift[kMaxFields][27];
};

template <int TB, int TZ>
__device__ __forceinline__ void nb_fill3(NbTable3& nb, CInst& I, int x0, int y0, int z0, bool regular,
                                         int nsets, int ncons) {
    const int d = threadIdx.x;
    if (d >= 27) return;
    const int dx = d % 3 - 1, dy = (d / 3) % 3 - 1, dz = d / 9 - 1;
    const int rx = dx < 0 ? x0 - 1 : (dx > 0 ? x0 + TB : x0);
    const int ry = dy < 0 ? y0 - 1 : (dy > 0 ? y0 + TB : y0);
    const int rz = dz < 0 ? z0 - 1 : (dz > 0 ? z0 + TZ : z0);
    const int mx = map_index(rx, I.nx, 0), my = map_index(ry, I.ny, 0), mz = map_index(rz, I.nz, 0);
    const int tt = mx / TB + (my / TB) * I.ntx + (mz / TZ) * I.ntx * I.nty;
    nb.tile[d] = tt;
    // both sets' bitmap words loaded together (a one-set instance reads set 0
    // twice), from pointers selected out of one scalar batch, so no load waits
    // on a branch or on a pointer load of its own
    const gu64 *par0 = I.par[0], *par1 = I.par[1], *cur0 = I.a_cur[0], *cur1 = I.a_cur[1];
    asm volatile("" : "+s"(par0), "+s"(par1), "+s"(cur0), "+s"(cur1));
    const bool two = nsets > 1;
    const uint64_t wp0 = par0[tt >> 6], wp1 = (two ? par1 : par0)[tt >> 6];
    const uint64_t wa0 = cur0[tt >> 6], wa1 = (two ? cur1 : cur0)[tt >> 6];
    const int b = tt & 63;
    nb.par[0][d] = (int)((wp0 >> b) & 1ull);
    nb.act[0][d] = regular ? 1 : (int)((wa0 >> b) & 1ull);
    if (two) {
        nb.par[1][d] = (int)((wp1 >> b) & 1ull);
        nb.act[1][d] = regular ? 1 : (int)((wa1 >> b) & 1ull);
    }
    const int p0 = (int)((wp0 >> b) & 1ull);
    for (int f = 0; f < kMaxFields; f++)
        nb.shift[f][d] = f < ncons ? (I.scal[8 + f] - I.s_tile[((int64_t)f * 2 + p0) * I.nt + tt]) : 0.0;
}

// 3D wrap of a small periodic offset
__device__ __forceinline__ int wrapi(int i, int n) {
    while (i < 0) i += n;
    while (i >= n) i -= n;
    return i;
}

// 27-point isotropic Laplacian from an LDS box with row stride RX and plane
// stride PZ, in the summation order of example_Reactive_ModelB/kernel.cu:333-344.
template <int RX, int PZ>
__device__ __forceinline__ double lds_lap27(const double* s, int x, int y, int z, double inv_dx2) {
#define S3(a, b, c) s[(x + (a)) + (y + (b)) * RX + (z + (c)) * PZ]
    return inv_dx2 *
           (-64.0 / 15.0 * S3(0, 0, 0) +
            7.0 / 15.0 * (S3(1, 0, 0) + S3(-1, 0, 0) + S3(0, 1, 0) + S3(0, -1, 0) + S3(0, 0, 1) + S3(0, 0, -1)) +
            0.1 * (S3(1, 1, 0) + S3(1, -1, 0) + S3(-1, 1, 0) + S3(-1, -1, 0) + S3(1, 0, 1) + S3(0, 1, 1) +
                   S3(0, -1, 1) + S3(-1, 0, 1) + S3(1, 0, -1) + S3(0, 1, -1) + S3(0, -1, -1) + S3(-1, 0, -1)) +
            1.0 / 30.0 * (S3(1, 1, 1) + S3(-1, 1, 1) + S3(-1, -1, 1) + S3(1, -1, 1) + S3(1, 1, -1) +
                          S3(-1, 1, -1) + S3(-1, -1, -1) + S3(1, -1, -1)));
#undef S3
}

template <int TB, int TZ, int H>
__device__ __forceinline__ int nb_dir3(int lx, int ly, int lz) {
    const int a = lx < H ? 0 : (lx < TB + H ? 1 : 2);
    const int b = ly < H ? 0 : (ly < TB + H ? 1 : 2);
    const int c = lz < H ? 0 : (lz < TZ + H ? 1 : 2);
    return a + 3 * b + 9 * c;
}

// ------------------------------------------------------------ Model B (ternary)
// Two launches per step, like the reference's Calc_mu / Update pair: the first
// computes the chemical potentials of every awake block's own cells into the
// stored-mu arrays (aux 0..2), the second updates x, y, z from the potentials
// of the block and its 1-cell ring. Sleeping neighbours keep the potentials they
// stored when they last ran, so the ring needs no recomputation and each cell's
// logarithms are taken once per step.
//
// One wave per 4x4x4 block, one cell per lane. Both kernels are software
// pipelined over the workgroup's blocks (resolved 64 rounds at a time by
// round_list): while block i computes from LDS, block i+1's box is in flight
// into registers and block i+2's parity words are in flight, so no block waits
// on a dependent chain of loads (worklist -> parities -> box).

// The halo box elements a lane loads (k = lane + 64 i), packed as
// lx | ly << 4 | lz << 8 | neighbour direction << 12 (the same for every block).
template <int TB, int TZ>
struct Box1 {
    static constexpr int WX = TB + 2, WZ = TZ + 2, A = WX * WX * WZ, NH = (A + 63) / 64;
    static constexpr int RX = WX, PZ = WX * WX, AP = A;  // LDS row and plane strides
    int c[NH];
    __device__ __forceinline__ void init(int lane) {
#pragma unroll
        for (int i = 0; i < NH; i++) {
            const int k = lane + i * 64;
            const int lx = k % WX, ly = (k / WX) % WX, lz = k / (WX * WX);
            c[i] = lx | ly << 4 | lz << 8 | nb_dir3<TB, TZ, 1>(lx, ly, lz) << 12;
        }
    }
};

// v in [-1, n] wrapped into [0, n)
__device__ __forceinline__ int wrap1(int v, int n) {
    v = v < 0 ? v + n : v;
    return v >= n ? v - n : v;
}

// Band sweep, for a single instance whose list holds every block (the usual
// case of a fully awake 3D field). Workgroups are dealt to the 8 XCDs
// round-robin by index, so XCD x = b mod 8 owns the band of Cx = ntx*nty/8
// consecutive block columns [x Cx, (x+1) Cx) and sweeps it z-layer by z-layer:
// the band's blocks in (z-block, column) order are dealt round-robin to the
// XCD's S = G/8 workgroups (item i = b/8 + k S). With a grid that is resident
// at once, the XCD's workgroups work on a contiguous run of that order at any
// moment, so a halo plane is still in the XCD's L2 when the blocks on its
// other side need it, and only the two band edges read across XCDs.
// cx = 0 where the grid or the list does not allow it (list order applies).
struct BandSweep {
    int cx, c, per, s, j, base;
};

__device__ __forceinline__ BandSweep band_sweep_setup(const WorkList& w, int total) {
    BandSweep bs{0, 0, 0, 0, 0, 0};
    // one instance (segments are in instance order) whose list is full
    if (w.nseg < 1 || w.segs[w.nseg - 1].inst != 0) return bs;
    const CInst& I = *(const CInst*)w.insts;
    if (total != I.nt) return bs;
    const int G = gridDim.x, C = I.ntx * I.nty;
    if ((G & 7) != 0 || (C & 7) != 0) return bs;
    bs.cx = C >> 3;
    bs.c = C;
    bs.per = I.nt >> 3;
    bs.s = G >> 3;
    bs.j = blockIdx.x >> 3;
    bs.base = (blockIdx.x & 7) * bs.cx;
    return bs;
}

// The sweep's blocks k = round0 .. round0+63 of this workgroup in the RoundList
// form (lane k - round0: item b/8 + k S of its XCD's order).
__device__ __forceinline__ RoundList band_sweep_round(const BandSweep& bs, int round0) {
    const int64_t i = bs.j + (int64_t)(round0 + (int)(threadIdx.x & 63)) * bs.s;
    const bool ok = i < bs.per;
    RoundList L;
    L.inst = 0;
    L.t = ok ? (int)(i / bs.cx) * bs.c + bs.base + (int)(i % bs.cx) : 0;
    L.valid = __ballot(ok);
    return L;
}

// A block of the worklist and what its loads need (all wave-uniform).
struct MbBlock {
    CInst* I;
    int t, x0, y0, z0;
    uint64_t pm;  // parity of the 27 neighbour blocks, bit d = (dx+1) + 3 (dy+1) + 9 (dz+1)
};

template <int TB, int TZ>
__device__ __forceinline__ MbBlock mb_take(const WorkList& w, const RoundList& L, uint64_t& todo) {
    MbBlock b;
    round_take(w, L, todo, &b.I, &b.t);
    CInst& I = *b.I;
    b.x0 = (b.t % I.ntx) * TB;
    b.y0 = ((b.t / I.ntx) % I.nty) * TB;
    b.z0 = (b.t / (I.ntx * I.nty)) * TZ;
    b.pm = 0;
    return b;
}

// Parity of neighbour block `lane` (lanes 0..26): the bitmap word is loaded
// here and the bit taken later (mb_parities_finish), so the load stays in flight.
struct ParityWord {
    uint64_t word;
    int tt;
};

template <int TB, int TZ>
__device__ __forceinline__ ParityWord mb_parities_issue(const MbBlock& b, int lane) {
    CInst& I = *b.I;
    ParityWord p{0ull, 0};
    if (lane < 27) {
        const int dx = lane % 3 - 1, dy = (lane / 3) % 3 - 1, dz = lane / 9 - 1;
        const int mx = wrap1(dx < 0 ? b.x0 - 1 : b.x0 + dx * TB, I.nx);
        const int my = wrap1(dy < 0 ? b.y0 - 1 : b.y0 + dy * TB, I.ny);
        const int mz = wrap1(dz < 0 ? b.z0 - 1 : b.z0 + dz * TZ, I.nz);
        p.tt = mx / TB + (my / TB) * I.ntx + (mz / TZ) * I.ntx * I.nty;
        p.word = I.par[0][p.tt >> 6];
    }
    return p;
}

// One ballot gathers the 27 bits; with conserved DBA lanes 0..26 also store each
// neighbour's pending shift S_f - epoch in sshift (read after a barrier).
__device__ __forceinline__ void mb_parities_finish(MbBlock& b, const ParityWord& p, int lane,
                                                   double (*sshift)[27]) {
    CInst& I = *b.I;
    const int bit = lane < 27 ? (int)((p.word >> (p.tt & 63)) & 1ull) : 0;
    b.pm = __ballot(bit);
    if (lane < 27)
        for (int f = 0; f < I.ncons; f++)
            sshift[f][lane] = I.scal[8 + f] - I.s_tile[((int64_t)f * 2 + bit) * I.nt + p.tt];
}

template <int TB, int TZ>
__device__ __forceinline__ int64_t mb_box_index(const MbBlock& b, int c) {
    CInst& I = *b.I;
    const int lx = c & 15, ly = (c >> 4) & 15, lz = (c >> 8) & 15;
    return wrap1(b.x0 + lx - 1, I.nx) + (int64_t)wrap1(b.y0 + ly - 1, I.ny) * I.nx +
           (int64_t)wrap1(b.z0 + lz - 1, I.nz) * ((int64_t)I.nx * I.ny);
}

// A block whose 1-cell box needs no periodic wrap addresses the box as one
// uniform base plus a per-lane offset (the same for every such block of the
// instance): no wrap selects and no 64-bit index arithmetic per element.
template <int TB, int TZ>
__device__ __forceinline__ bool mb_interior(const MbBlock& b) {
    CInst& I = *b.I;
    return b.x0 > 0 && b.x0 + TB < I.nx && b.y0 > 0 && b.y0 + TB < I.ny && b.z0 > 0 && b.z0 + TZ < I.nz &&
           (int64_t)I.nx * I.ny < (1 << 28);
}

// uniform index of the box corner (x0-1, y0-1, z0-1) of an interior block
__device__ __forceinline__ int64_t mb_box_base(const MbBlock& b) {
    CInst& I = *b.I;
    return (b.x0 - 1) + (int64_t)(b.y0 - 1) * I.nx + (int64_t)(b.z0 - 1) * ((int64_t)I.nx * I.ny);
}

// per-lane offset of packed box element c from the box corner
__device__ __forceinline__ int mb_box_offset(int c, int nx, int nxy) {
    return (c & 15) + ((c >> 4) & 15) * nx + ((c >> 8) & 15) * nxy;
}

template <int TB, int TZ>
__device__ __forceinline__ int64_t mb_own_index(const MbBlock& b, int lane) {
    CInst& I = *b.I;
    const int lx = lane % TB, ly = (lane / TB) % TB, lz = lane / (TB * TB);
    return (b.x0 + lx) + (int64_t)(b.y0 + ly) * I.nx + (int64_t)(b.z0 + lz) * ((int64_t)I.nx * I.ny);
}

// The chemical potentials of the cell at (cx, cy, cz) of an LDS box of x, y, z
// (field f at sc + f AP) into df; returns lap z for the detector
// (example_Reactive_ModelB/kernel.cu:557-587).
template <int RX, int PZ, int AP>
__device__ __forceinline__ double mb_mu_at(CInst& I, const double* sc, int cx, int cy, int cz, double phi,
                                           double (&df)[3]) {
    const int q = cx + cy * RX + cz * PZ;
    const double X = sc[q], Y = sc[AP + q], Z = sc[2 * AP + q];
    const double B = 1.0 - X - Y - Z;
    // logarithms first, then one Laplacian at a time: left free, the
    // scheduler hoists all 81 LDS reads above the logarithms (250 VGPRs)
    const double logB = log_f64(B), logX = log_f64(X), logY = log_f64(Y), logZ = log_f64(Z);
    __builtin_amdgcn_sched_barrier(0);
    const double inv_dx2 = I.invdx2;
    const double lx_ = lds_lap27<RX, PZ>(sc, cx, cy, cz, inv_dx2);
    __builtin_amdgcn_sched_barrier(0);
    const double ly_ = lds_lap27<RX, PZ>(sc + AP, cx, cy, cz, inv_dx2);
    __builtin_amdgcn_sched_barrier(0);
    const double lz_ = lds_lap27<RX, PZ>(sc + 2 * AP, cx, cy, cz, inv_dx2);
    df[0] = I.mb_cx + I.chi_xy * Y + I.chi_xz * Z + I.chi_xb * B - I.chi_xb * X - I.chi_yb * Y - I.chi_zb * Z +
            logX / I.r_x - logB - I.epsx_sq * lx_;
    df[1] = I.mb_cy + I.chi_xy * X + I.chi_yz * Z + I.chi_yb * B - I.chi_xb * X - I.chi_yb * Y - I.chi_zb * Z +
            logY / I.r_y - logB - I.epsy_sq * ly_;
    df[2] = I.mb_cz + I.chi_xz * X + I.chi_yz * Y + I.chi_zb * B - I.chi_xb * X - I.chi_yb * Y - I.chi_zb * Z +
            logZ / I.r_z - logB - I.p_gel * phi * phi / 2.0 / I.mb_omz - I.epsz_sq * lz_ - I.mb_lk;
    return lz_;
}

// The changes of x, y, z at (cx, cy, cz) from an LDS box of the potentials
// (example_Reactive_ModelB/kernel.cu:419-432): dt v_f (mob_f lap mu_f -/+ the reaction)
template <int RX, int PZ, int AP>
__device__ __forceinline__ void mb_change_at(CInst& I, const double* smu, int cx, int cy, int cz,
                                             double (&ch)[3]) {
    const double inv_dx2 = I.invdx2;
    const int mq = cx + cy * RX + cz * PZ;
    const double dfx = smu[mq], dfy = smu[AP + mq], dfz = smu[2 * AP + mq];
    const double R = I.k_0 * (exp(I.n_exp * I.v_x * dfx + I.m_exp * I.v_y * dfy) - exp(I.v_z * dfz));
    __builtin_amdgcn_sched_barrier(0);
    ch[0] = I.dt * I.v_x * (I.mob_x * lds_lap27<RX, PZ>(smu, cx, cy, cz, inv_dx2) - I.n_exp * R);
    __builtin_amdgcn_sched_barrier(0);
    ch[1] = I.dt * I.v_y * (I.mob_y * lds_lap27<RX, PZ>(smu + AP, cx, cy, cz, inv_dx2) - I.m_exp * R);
    __builtin_amdgcn_sched_barrier(0);
    ch[2] = I.dt * I.v_z * (I.mob_z * lds_lap27<RX, PZ>(smu + 2 * AP, cx, cy, cz, inv_dx2) + R);
}

// Calc_mu(_wakeup) + the detector |lap z| > eps (example_Reactive_ModelB/kernel.cu:290-420)
template <int TB, int TZ>
__global__ __launch_bounds__(64) void modelb_mu_kernel(const WorkList wl, int sb_from) {
    using B1 = Box1<TB, TZ>;
    constexpr int A = B1::A, NH = B1::NH, C = TB * TB * TZ, RX = B1::RX, PZ = B1::PZ, AP = B1::AP;
    static_assert(C <= 64, "at most one cell per lane");
    __shared__ double sc[3 * AP];
    __shared__ double sshift[2][kMaxFields][27];

    const int lane = threadIdx.x;
    B1 box;
    box.init(lane);
    double h[NH][3];
    // block q's box (parities known) into h, its own phi into ph
    constexpr uint64_t kAll27 = (1ull << 27) - 1;
    auto load = [&](const MbBlock& q, double& ph) {
        CInst& I = *q.I;
        if (mb_interior<TB, TZ>(q) && (q.pm == 0 || q.pm == kAll27)) {
            // no wrap, and the 27 blocks current in one buffer
            const int p = (int)(q.pm & 1);
            const int64_t base = mb_box_base(q);
            const int nx = I.nx, nxy = I.nx * I.ny;
            const double *b0 = I.buf[p][0] + base, *b1 = I.buf[p][1] + base, *b2 = I.buf[p][2] + base;
#pragma unroll
            for (int i = 0; i < NH; i++) {
                if (lane + i * 64 < A) {
                    const int o = mb_box_offset(box.c[i], nx, nxy);
                    h[i][0] = b0[o];
                    h[i][1] = b1[o];
                    h[i][2] = b2[o];
                }
            }
        } else {
#pragma unroll
            for (int i = 0; i < NH; i++) {
                if (lane + i * 64 < A) {
                    const int c = box.c[i];
                    const int64_t gi = mb_box_index<TB, TZ>(q, c);
                    const int pp = (int)(q.pm >> (c >> 12)) & 1;
                    for (int f = 0; f < 3; f++) h[i][f] = I.buf[pp][f][gi];
                }
            }
        }
        if (lane < C) ph = I.buf[(int)(q.pm >> 13) & 1][3][mb_own_index<TB, TZ>(q, lane)];
    };

    const int total = wl.offsets[wl.nseg];
    // a list of at least sb_from blocks is the superblock kernels' step
    if (total >= sb_from) return;
    const BandSweep bs = band_sweep_setup(wl, total);
    for (int round0 = 0; (int64_t)round0 * gridDim.x < total; round0 += 64) {
        const RoundList L = bs.cx ? band_sweep_round(bs, round0) : round_list(wl, total, round0);
        uint64_t todo = L.valid;
        if (!todo) continue;
        MbBlock b = mb_take<TB, TZ>(wl, L, todo);
        mb_parities_finish(b, mb_parities_issue<TB, TZ>(b, lane), lane, sshift[0]);
        __syncthreads();
        double phi = 0.0;
        load(b, phi);
        bool has1 = todo != 0;
        MbBlock n1{};
        ParityWord w1{0ull, 0};
        if (has1) {
            n1 = mb_take<TB, TZ>(wl, L, todo);
            w1 = mb_parities_issue<TB, TZ>(n1, lane);
        }
        int slot = 0;
        for (;;) {
            CInst& I = *b.I;
            const int ncons = I.ncons;
            // block b's box into LDS, each value in the current frame
#pragma unroll
            for (int i = 0; i < NH; i++) {
                const int k = lane + i * 64;
                if (k < A) {
                    const int d = box.c[i] >> 12;
                    for (int f = 0; f < 3; f++) {
                        double v = h[i][f];
                        if (f < ncons) v -= sshift[slot][f][d];
                        sc[f * AP + k] = v;
                    }
                }
            }
            __syncthreads();
            double phi_n = 0.0;
            bool has2 = false;
            MbBlock n2{};
            ParityWord w2{0ull, 0};
            if (has1) {
                mb_parities_finish(n1, w1, lane, sshift[slot ^ 1]);
                load(n1, phi_n);
                if (todo) {
                    n2 = mb_take<TB, TZ>(wl, L, todo);
                    w2 = mb_parities_issue<TB, TZ>(n2, lane);
                    has2 = true;
                }
            }

            int flag = 0;
            if (lane < C) {
                const int lx = lane % TB, ly = (lane / TB) % TB, lz = lane / (TB * TB);
                double df[3];
                const double lz_ = mb_mu_at<RX, PZ, AP>(I, sc, lx + 1, ly + 1, lz + 1, phi, df);
                const int64_t gi = mb_own_index<TB, TZ>(b, lane);
                I.aux[0][gi] = df[0];
                I.aux[1][gi] = df[1];
                I.aux[2][gi] = df[2];
                flag = fabs(lz_) > I.thresh_a;
            }
            if (I.mode != 0) {
                const int any = __ballot(flag) != 0ull;
                if (lane == 0) I.nflag[0][b.t] = (uint8_t)any;
            }
            __syncthreads();
            if (!has1) break;
            b = n1;
            phi = phi_n;
            n1 = n2;
            w1 = w2;
            has1 = has2;
            slot ^= 1;
        }
    }
}

// Update(_wakeup) + copy-back + conserved change capture
// (example_Reactive_ModelB/kernel.cu:430-560, Conserved/kernel.cu:630-702)
template <int TB, int TZ>
__global__ __launch_bounds__(64) void modelb_update_kernel(const WorkList wl, int sb_from) {
    using B1 = Box1<TB, TZ>;
    constexpr int A = B1::A, NH = B1::NH, C = TB * TB * TZ, RX = B1::RX, PZ = B1::PZ, AP = B1::AP;
    static_assert(C <= 64, "at most one cell per lane");
    __shared__ double smu[3 * AP];

    const int lane = threadIdx.x;
    B1 box;
    box.init(lane);
    double h[NH][3];
    // block q's potentials box into h, its own x, y, z (buffer pq) into vq and
    // their shift epochs into eq (conserved)
    auto load = [&](const MbBlock& q, int pq, double (&vq)[3], double (&eq)[3]) {
        CInst& I = *q.I;
        if (mb_interior<TB, TZ>(q)) {
            const int64_t base = mb_box_base(q);
            const int nx = I.nx, nxy = I.nx * I.ny;
            const double *a0 = I.aux[0] + base, *a1 = I.aux[1] + base, *a2 = I.aux[2] + base;
#pragma unroll
            for (int i = 0; i < NH; i++) {
                if (lane + i * 64 < A) {
                    const int o = mb_box_offset(box.c[i], nx, nxy);
                    h[i][0] = a0[o];
                    h[i][1] = a1[o];
                    h[i][2] = a2[o];
                }
            }
        } else {
#pragma unroll
            for (int i = 0; i < NH; i++) {
                if (lane + i * 64 < A) {
                    const int64_t gi = mb_box_index<TB, TZ>(q, box.c[i]);
                    for (int f = 0; f < 3; f++) h[i][f] = I.aux[f][gi];
                }
            }
        }
        if (lane < C) {
            const int64_t gi = mb_own_index<TB, TZ>(q, lane);
            for (int f = 0; f < 3; f++) vq[f] = I.buf[pq][f][gi];
        }
        for (int f = 0; f < I.ncons; f++) eq[f] = I.s_tile[((int64_t)f * 2 + pq) * I.nt + q.t];
    };

    const int total = wl.offsets[wl.nseg];
    // a list of at least sb_from blocks is the superblock kernels' step
    if (total >= sb_from) return;
    const BandSweep bs = band_sweep_setup(wl, total);
    for (int round0 = 0; (int64_t)round0 * gridDim.x < total; round0 += 64) {
        const RoundList L = bs.cx ? band_sweep_round(bs, round0) : round_list(wl, total, round0);
        uint64_t todo = L.valid;
        if (!todo) continue;
        MbBlock b = mb_take<TB, TZ>(wl, L, todo);
        int pself = bit_of(b.I->par[0], b.t);
        double v[3] = {0.0, 0.0, 0.0}, e[3] = {0.0, 0.0, 0.0};
        load(b, pself, v, e);
        bool has1 = todo != 0;
        MbBlock n1{};
        uint64_t w1 = 0;
        if (has1) {
            n1 = mb_take<TB, TZ>(wl, L, todo);
            w1 = n1.I->par[0][n1.t >> 6];
        }
        for (;;) {
#pragma unroll
            for (int i = 0; i < NH; i++) {
                const int k = lane + i * 64;
                if (k < A)
                    for (int f = 0; f < 3; f++) smu[f * AP + k] = h[i][f];
            }
            __syncthreads();
            double vn[3] = {0.0, 0.0, 0.0}, en[3] = {0.0, 0.0, 0.0};
            int pn = 0;
            bool has2 = false;
            MbBlock n2{};
            uint64_t w2 = 0;
            if (has1) {
                pn = (int)((w1 >> (n1.t & 63)) & 1ull);
                load(n1, pn, vn, en);
                if (todo) {
                    n2 = mb_take<TB, TZ>(wl, L, todo);
                    w2 = n2.I->par[0][n2.t >> 6];
                    has2 = true;
                }
            }

            CInst& I = *b.I;
            const int ncons = I.ncons;
            double part[3] = {0.0, 0.0, 0.0};
            if (lane < C) {
                const int lx = lane % TB, ly = (lane / TB) % TB, lz = lane / (TB * TB);
                double x[3] = {v[0], v[1], v[2]};
                for (int f = 0; f < ncons; f++) x[f] -= I.scal[8 + f] - e[f];
                double ch[3];
                mb_change_at<RX, PZ, AP>(I, smu, lx + 1, ly + 1, lz + 1, ch);
                const int64_t gi = mb_own_index<TB, TZ>(b, lane);
                I.buf[pself ^ 1][0][gi] = x[0] + ch[0];
                I.buf[pself ^ 1][1][gi] = x[1] + ch[1];
                I.buf[pself ^ 1][2][gi] = x[2] + ch[2];
                part[0] = ch[0];
                part[1] = ch[1];
                part[2] = ch[2];
            }
            for (int f = 0; f < ncons; f++) {
                const double tot = wave_sum_down(part[f]);
                if (lane == 0) {
                    I.tpart[(int64_t)f * I.nt + b.t] = tot;
                    I.s_tile[((int64_t)f * 2 + (pself ^ 1)) * I.nt + b.t] = I.scal[8 + f];
                }
            }
            __syncthreads();
            if (!has1) break;
            b = n1;
            pself = pn;
            for (int f = 0; f < 3; f++) {
                v[f] = vn[f];
                e[f] = en[f];
            }
            n1 = n2;
            w1 = w2;
            has1 = has2;
        }
    }
}

// ------------------------------------------------------------ Model B, superblocks
// Steps of one 3D Model B instance (4x4x4 blocks, no conserved shift) in which
// at least sb_from blocks are awake run on superblocks of 2 x 2 x SZ blocks: one
// workgroup of 4 SZ waves stages the box of a superblock's 8 x 8 x 4SZ cells and
// their 1-cell halo once (10 x 10 x (4SZ + 2) values per field; each block's own
// box holds 216 values for 64 cells), and wave w computes block w of the
// superblock if that block is awake. The arithmetic per cell is mb_mu_at /
// mb_change_at, as in the block kernels, so both give the same doubles. Each
// superblock kernel returns at once when the list is shorter than sb_from, and
// the block kernels when it is not.

// positions of the blocks around a superblock among the 4 x 4 x 4 blocks its box
// can touch: bit (ox + 1) + 4 (oy + 1) + 16 (oz + 1), block offsets o in [-1, 2]
__device__ __forceinline__ int sb_code(int w) { return 21 + (w & 1) + 4 * ((w >> 1) & 1) + 16 * (w >> 2); }

template <int SZ>
struct Sb {
    static constexpr int W = 10, WZ = 4 * SZ + 2, A = W * W * WZ;  // box sides and size
    static constexpr int NT = 256 * SZ;                            // 4 SZ waves
    static constexpr int NE = (A + NT - 1) / NT;                   // box elements per thread
    static constexpr uint64_t own() {
        uint64_t m = 0;
        for (int w = 0; w < 4 * SZ; w++) m |= 1ull << (21 + (w & 1) + 4 * ((w >> 1) & 1) + 16 * (w >> 2));
        return m;
    }
};

// A superblock as every wave of its workgroup sees it (all uniform).
struct MbSuper {
    int X0, Y0, Z0;  // first cell
    int tw;          // this wave's block
    uint64_t pm;     // parity of the blocks the box touches (bit 1: buf[1] is current)
    uint64_t am;     // awake bits of the own blocks, at the same positions
    bool interior;   // the box needs no periodic wrap, and its offsets fit an int
};

// Each wave reads the parity and awake words itself (lane l: block l of the
// 4 x 4 x 4 around the superblock's corner block), so no barrier is needed
// before the box loads.
template <int SZ>
__device__ __forceinline__ MbSuper sb_setup(CInst& I, int sb, bool regular) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int ntx = I.ntx, nty = I.nty, nsx = ntx >> 1, nsy = nty >> 1;
    const int sx = sb % nsx, sy = (sb / nsx) % nsy, sz = sb / (nsx * nsy);
    MbSuper s;
    s.X0 = sx * 8;
    s.Y0 = sy * 8;
    s.Z0 = sz * 4 * SZ;
    s.tw = (2 * sx + (w & 1)) + (2 * sy + ((w >> 1) & 1)) * ntx + (SZ * sz + (w >> 2)) * ntx * nty;
    // a one-block-thick superblock's box reaches oz = 1 at most (lanes past it read that layer)
    const int oz = min(lane >> 4, SZ + 1) - 1;
    const int bx = wrap1(2 * sx + (lane & 3) - 1, ntx), by = wrap1(2 * sy + ((lane >> 2) & 3) - 1, nty),
              bz = wrap1(SZ * sz + oz, I.ntz);
    const int tt = bx + by * ntx + bz * ntx * nty;
    const gu64 *par0 = I.par[0], *cur0 = I.a_cur[0];
    asm volatile("" : "+s"(par0), "+s"(cur0));
    const uint64_t wp = par0[tt >> 6], wa = cur0[tt >> 6];
    constexpr uint64_t kOwn = Sb<SZ>::own();
    s.pm = __ballot((int)((wp >> (tt & 63)) & 1ull));
    s.am = regular ? kOwn : (__ballot((int)((wa >> (tt & 63)) & 1ull)) & kOwn);
    s.interior = s.X0 > 0 && s.X0 + 8 < I.nx && s.Y0 > 0 && s.Y0 + 8 < I.ny && s.Z0 > 0 &&
                 s.Z0 + 4 * SZ < I.nz && (int64_t)I.nx * I.ny < (1 << 27);
    return s;
}

// The superblocks of this workgroup, k = 0, 1, ...: the band sweep of the block
// kernels over superblock columns (XCD b mod 8 sweeps its band of columns
// z-layer by z-layer), or plain grid-stride order where the counts do not allow it.
template <int SZ>
__device__ __forceinline__ int sb_of(CInst& I, int k) {
    const int G = gridDim.x, nsb = I.nt / (4 * SZ), C = (I.ntx >> 1) * (I.nty >> 1);
    if ((G & 7) == 0 && (C & 7) == 0) {
        const int cx = C >> 3;
        const int64_t i = (int64_t)(blockIdx.x >> 3) + (int64_t)k * (G >> 3);
        if (i >= (nsb >> 3)) return -1;
        return (int)(i / cx) * C + (int)(blockIdx.x & 7) * cx + (int)(i % cx);
    }
    const int64_t i = (int64_t)blockIdx.x + (int64_t)k * G;
    return i < nsb ? (int)i : -1;
}

// The thread's box elements k = tid + e NT (the last round's slots past the box
// hold element tid's coordinates and are neither loaded nor stored).
template <int SZ>
struct SbElems {
    using S = Sb<SZ>;
    int lx[S::NE], ly[S::NE], lz[S::NE], d[S::NE], off[S::NE];
    __device__ __forceinline__ void init(int tid, int nx, int nxy) {
#pragma unroll
        for (int e = 0; e < S::NE; e++) {
            const int k0 = tid + e * S::NT, k = k0 < S::A ? k0 : tid;
            lx[e] = k % S::W;
            ly[e] = (k / S::W) % S::W;
            lz[e] = k / (S::W * S::W);
            d[e] = ((lx[e] + 3) >> 2) + 4 * ((ly[e] + 3) >> 2) + 16 * ((lz[e] + 3) >> 2);
            off[e] = lx[e] + ly[e] * nx + lz[e] * nxy;
        }
    }
    // wrapped global index of element e of the box at (X0, Y0, Z0)
    __device__ __forceinline__ int64_t index(CInst& I, const MbSuper& s, int e) const {
        return wrap1(s.X0 + lx[e] - 1, I.nx) + (int64_t)wrap1(s.Y0 + ly[e] - 1, I.ny) * I.nx +
               (int64_t)wrap1(s.Z0 + lz[e] - 1, I.nz) * ((int64_t)I.nx * I.ny);
    }
};

// Calc_mu(_wakeup) + the detector on superblocks (see modelb_mu_kernel)
// 4 waves per SIMD (128 VGPRs; left free the compiler takes 129 and 3 waves):
// measured 0.945 ms per 256^3 step with the update kernel at 5, 1.04 ms at 3
template <int SZ>
__global__ __launch_bounds__(Sb<SZ>::NT) __attribute__((amdgpu_waves_per_eu(4))) void modelb_mu_sb_kernel(const WorkList wl, int sb_from) {
    using S = Sb<SZ>;
    constexpr int A = S::A, NT = S::NT, NE = S::NE;
    __shared__ double sc[3 * A];
    const int total = wl.offsets[wl.nseg];
    if (total < sb_from) return;
    CInst& I = *(CInst*)wl.insts;
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const bool regular = I.mode == 0;
    const int nx = I.nx, nxy = I.nx * I.ny;
    SbElems<SZ> el;
    el.init(tid, nx, nxy);
    const int cx = 4 * (w & 1) + (lane & 3), cy = 4 * ((w >> 1) & 1) + ((lane >> 2) & 3), cz = 4 * (w >> 2) + (lane >> 4);
    const int code = sb_code(w);
    for (int k = 0;; k++) {
        const int sb = sb_of<SZ>(I, k);
        if (sb < 0) break;
        const MbSuper s = sb_setup<SZ>(I, sb, regular);
        if (!s.am) continue;
        const bool awake = (s.am >> code) & 1ull;
        const int pw = (int)(s.pm >> code) & 1;
        const int64_t own = (s.X0 + cx) + (int64_t)(s.Y0 + cy) * nx + (int64_t)(s.Z0 + cz) * nxy;
        double h[NE][3];
        if (s.interior && (s.pm == 0ull || s.pm == ~0ull)) {
            // no wrap, and every block the box touches current in one buffer
            const int p = (int)(s.pm & 1ull);
            const int64_t base = (s.X0 - 1) + (int64_t)(s.Y0 - 1) * nx + (int64_t)(s.Z0 - 1) * nxy;
            const gdouble *b0 = I.buf[p][0] + base, *b1 = I.buf[p][1] + base, *b2 = I.buf[p][2] + base;
#pragma unroll
            for (int e = 0; e < NE; e++) {
                if (e + 1 < NE || tid + e * NT < A) {
                    h[e][0] = b0[el.off[e]];
                    h[e][1] = b1[el.off[e]];
                    h[e][2] = b2[el.off[e]];
                }
            }
        } else {
#pragma unroll
            for (int e = 0; e < NE; e++) {
                if (e + 1 < NE || tid + e * NT < A) {
                    const int64_t gi = el.index(I, s, e);
                    const int pp = (int)(s.pm >> el.d[e]) & 1;
                    for (int f = 0; f < 3; f++) h[e][f] = I.buf[pp][f][gi];
                }
            }
        }
        double phi = 0.0;
        if (awake) phi = I.buf[pw][3][own];
#pragma unroll
        for (int e = 0; e < NE; e++)
            if (tid + e * NT < A)
                for (int f = 0; f < 3; f++) sc[f * A + tid + e * NT] = h[e][f];
        __syncthreads();
        if (awake) {
            double df[3];
            const double lz_ = mb_mu_at<S::W, S::W * S::W, A>(I, sc, cx + 1, cy + 1, cz + 1, phi, df);
            I.aux[0][own] = df[0];
            I.aux[1][own] = df[1];
            I.aux[2][own] = df[2];
            if (I.mode != 0) {
                const int any = __ballot(fabs(lz_) > I.thresh_a) != 0ull;
                if (lane == 0) I.nflag[0][s.tw] = (uint8_t)any;
            }
        }
        __syncthreads();
    }
}

// Update(_wakeup) + copy-back on superblocks (see modelb_update_kernel)
// 5 waves per SIMD (96 VGPRs, no spills; 4 left free)
template <int SZ>
__global__ __launch_bounds__(Sb<SZ>::NT) __attribute__((amdgpu_waves_per_eu(5))) void modelb_update_sb_kernel(const WorkList wl, int sb_from) {
    using S = Sb<SZ>;
    constexpr int A = S::A, NT = S::NT, NE = S::NE;
    __shared__ double smu[3 * A];
    const int total = wl.offsets[wl.nseg];
    if (total < sb_from) return;
    CInst& I = *(CInst*)wl.insts;
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const bool regular = I.mode == 0;
    const int nx = I.nx, nxy = I.nx * I.ny;
    SbElems<SZ> el;
    el.init(tid, nx, nxy);
    const int cx = 4 * (w & 1) + (lane & 3), cy = 4 * ((w >> 1) & 1) + ((lane >> 2) & 3), cz = 4 * (w >> 2) + (lane >> 4);
    const int code = sb_code(w);
    for (int k = 0;; k++) {
        const int sb = sb_of<SZ>(I, k);
        if (sb < 0) break;
        const MbSuper s = sb_setup<SZ>(I, sb, regular);
        if (!s.am) continue;
        const bool awake = (s.am >> code) & 1ull;
        const int pw = (int)(s.pm >> code) & 1;
        const int64_t own = (s.X0 + cx) + (int64_t)(s.Y0 + cy) * nx + (int64_t)(s.Z0 + cz) * nxy;
        double h[NE][3];
        if (s.interior) {
            const int64_t base = (s.X0 - 1) + (int64_t)(s.Y0 - 1) * nx + (int64_t)(s.Z0 - 1) * nxy;
            const gdouble *a0 = I.aux[0] + base, *a1 = I.aux[1] + base, *a2 = I.aux[2] + base;
#pragma unroll
            for (int e = 0; e < NE; e++) {
                if (e + 1 < NE || tid + e * NT < A) {
                    h[e][0] = a0[el.off[e]];
                    h[e][1] = a1[el.off[e]];
                    h[e][2] = a2[el.off[e]];
                }
            }
        } else {
#pragma unroll
            for (int e = 0; e < NE; e++) {
                if (e + 1 < NE || tid + e * NT < A) {
                    const int64_t gi = el.index(I, s, e);
                    for (int f = 0; f < 3; f++) h[e][f] = I.aux[f][gi];
                }
            }
        }
        double v[3] = {0.0, 0.0, 0.0};
        if (awake)
            for (int f = 0; f < 3; f++) v[f] = I.buf[pw][f][own];
#pragma unroll
        for (int e = 0; e < NE; e++)
            if (tid + e * NT < A)
                for (int f = 0; f < 3; f++) smu[f * A + tid + e * NT] = h[e][f];
        __syncthreads();
        if (awake) {
            double ch[3];
            mb_change_at<S::W, S::W * S::W, A>(I, smu, cx + 1, cy + 1, cz + 1, ch);
            I.buf[pw ^ 1][0][own] = v[0] + ch[0];
            I.buf[pw ^ 1][1][own] = v[1] + ch[1];
            I.buf[pw ^ 1][2][own] = v[2] + ch[2];
        }
        __syncthreads();
    }
}

// ------------------------------------------------------------ Kobayashi-94 dendrite
template <int TB, int TZ>
__global__ __launch_bounds__(64) void k94_step_kernel(const WorkList wl) {
    constexpr int NT = 64;
    constexpr int WX = TB + 2, WZ = TZ + 2;
    constexpr int A = WX * WX * WZ;
    __shared__ double sp[A];
    __shared__ double sT[A];
    __shared__ int sflag[2];
    __shared__ NbTable3 nb;

    WorkCursor cur;
    work_begin(wl, cur, false);
    const int total = wl.offsets[wl.nseg];
    const int tid = threadIdx.x;
    for (;;) {
        CInst* Ip;
        int t;
        uint64_t rec;
        if (!work_next(wl, cur, total, &Ip, &t, &rec, false)) break;
        CInst& I = *Ip;
        const int nx = I.nx, ny = I.ny, nz = I.nz;
        const int x0 = (t % I.ntx) * TB, y0 = ((t / I.ntx) % I.nty) * TB, z0 = (t / (I.ntx * I.nty)) * TZ;
        const bool regular = I.mode == 0;
        const int64_t nxy = (int64_t)nx * ny;
        nb_fill3<TB, TZ>(nb, I, x0, y0, z0, regular, 2, 0);
        __syncthreads();
        const int ap = nb.act[0][13], at = nb.act[1][13];
        const int pp0 = nb.par[0][13], pp1 = nb.par[1][13];
        // stage phi and T of the block and the face cells of its 1-cell halo (the
        // 7-point stencils read no edge or corner of the box: 160 of its 216 cells
        // at 4x4x4): every load issued before the first wait (a slot past the cells
        // loads cell 0 of the list and is not stored), from buffer pointers of one
        // scalar batch. The cells: the block's column with its two z faces, then the
        // two x faces, then the two y faces.
        {
            constexpr int NCOL = TB * TB * (TZ + 2), NF = TB * TZ, NC = NCOL + 4 * NF;
            constexpr int KA = (NC + NT - 1) / NT;
            const gdouble *bp0 = I.buf[0][0], *bp1 = I.buf[1][0], *bT0 = I.buf[0][1], *bT1 = I.buf[1][1];
            asm volatile("" : "+s"(bp0), "+s"(bp1), "+s"(bT0), "+s"(bT1));
            double vp[KA], vT[KA];
            int dk[KA], kb[KA];
#pragma unroll
            for (int q = 0; q < KA; q++) {
                const int k0 = tid + q * NT, k = k0 < NC ? k0 : 0;
                int lx, ly, lz;
                if (k < NCOL) {
                    lx = 1 + k % TB;
                    ly = 1 + (k / TB) % TB;
                    lz = k / (TB * TB);
                } else {
                    const int j = k - NCOL, face = j / NF, jj = j % NF;
                    const int u = 1 + jj % TB, edge = (face & 1) ? TB + 1 : 0;
                    lz = 1 + jj / TB;
                    lx = face < 2 ? edge : u;
                    ly = face < 2 ? u : edge;
                }
                const int d = nb_dir3<TB, TZ, 1>(lx, ly, lz);
                const int64_t gi = wrapi(x0 + lx - 1, nx) + wrapi(y0 + ly - 1, ny) * (int64_t)nx +
                                   wrapi(z0 + lz - 1, nz) * nxy;
                vp[q] = (nb.par[0][d] ? bp1 : bp0)[gi];
                vT[q] = (nb.par[1][d] ? bT1 : bT0)[gi];
                dk[q] = d;
                kb[q] = lx + ly * WX + lz * WX * WX;
            }
#pragma unroll
            for (int q = 0; q < KA; q++) {
                if (tid + q * NT < NC) {
                    if (dk[q] == 13 || ap) sp[kb[q]] = vp[q];
                    if (dk[q] == 13 || at || regular) sT[kb[q]] = vT[q];
                }
            }
        }
        __syncthreads();
        const double f0 = 1. / (2 * I.dx), f1 = 1. / I.dx / I.dx, f2 = -I.alpha / 3.14159, f3 = 1. / I.tau;
        const double dt = I.dt, eps = I.eps_bar, delta = I.delta, gam = I.gamma;
        int fp = 0, fT = 0;
        for (int k = tid; k < TB * TB * TZ; k += NT) {
            const int lx = k % TB, ly = (k / TB) % TB, lz = k / (TB * TB);
            const int64_t gi = (x0 + lx) + (int64_t)(y0 + ly) * nx + (int64_t)(z0 + lz) * nxy;
            const int q = (lx + 1) + (ly + 1) * WX + (lz + 1) * WX * WX;
            const double pc = sp[q], Tc = sT[q];
            double pn = pc;
            if (ap) {
                const double pxp = sp[q + 1], pxm = sp[q - 1], pyp = sp[q + WX], pym = sp[q - WX];
                const double pzp = sp[q + WX * WX], pzm = sp[q - WX * WX];
                const double dpx = (pxp - pxm) * f0, dpy = (pyp - pym) * f0, dpz = (pzp - pzm) * f0;
                const double lapP = (pxp + pxm + pyp + pym + pzp + pzm - 6 * pc) * f1;
                const double v2 = dpx * dpx + dpy * dpy + dpz * dpz;
                double sigma;
                if (v2 <= 1E-6)
                    sigma = 1;
                else
                    sigma = 1 - delta * (1 - ((dpx * dpx) * (dpx * dpx) + (dpy * dpy) * (dpy * dpy) +
                                              (dpz * dpz) * (dpz * dpz)) / (v2 * v2));
                const double m = f2 * atan(gam * sigma * Tc);
                const double p_change = f3 * (eps * eps * lapP + pc * (1 - pc) * (pc - 0.5 + m));
                pn = pc + dt * p_change;
                I.buf[pp0 ^ 1][0][gi] = pn;
                fp |= (pn >= I.thresh_a && pn <= (1.0 - I.thresh_a));
            }
            if (at) {
                const double lapT = (sT[q + 1] + sT[q - 1] + sT[q + WX] + sT[q - WX] + sT[q + WX * WX] +
                                     sT[q - WX * WX] - 6 * Tc) * f1;
                double tn;
                if (regular) {
                    tn = Tc + dt * lapT + pn - pc;
                } else {
                    tn = Tc + dt * lapT;
                    if (ap) tn += pn - pc;
                }
                I.buf[pp1 ^ 1][1][gi] = tn;
                fT |= fabs(tn - Tc) >= I.thresh_b;
            }
        }
        if (!regular) {
            const int anyp = block_any<NT>(fp, sflag);
            const int anyT = block_any<NT>(fT, sflag);
            if (tid == 0) {
                if (ap) I.nflag[0][t] = (uint8_t)anyp;
                if (at) I.nflag[1][t] = (uint8_t)anyT;
            }
        }
        __syncthreads();
    }
}

}  // namespace dba
