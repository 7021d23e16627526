// k_diag.h - diagnostics computed on the device from the current state, so a
// caller gets integrals, the dendrite growth front and the KH growth-mode
// amplitude without downloading whole fields.
//
// Reference counterparts: Integral (example_Reactive_ModelB/kernel.cu:989-1009,
// example_Kelvin_Helmholtz/KH_DBA.cu:1273-1288), growth_front / tip_curvature
// (example_Dendritic_growth/kernel.cu:760-790) and Calc_Vy_mode
// (example_Kelvin_Helmholtz/KH_DBA.cu:1291-1326), which run on the host over
// the whole managed arrays.
//
// Sums are deterministic: a fixed grid of kDiagBlocks x kDiagNT threads, each
// thread summing a fixed strided subset in index order, a fixed-order block
// reduction, then one thread adding the block partials in block order.
#pragma once
#include "common.h"

namespace dba {

constexpr int kDiagBlocks = 256;
constexpr int kDiagNT = 256;
constexpr int kMaxFrontCols = 64;

// part[b] = sum of field over the cells of block b.
__global__ __launch_bounds__(kDiagNT) void field_sum_kernel(const Inst* __restrict__ insts, int slot, int field,
                                                            double* __restrict__ part) {
    __shared__ double sred[kDiagNT / 64 + 1];
    CInst& I = ((CInst*)insts)[slot];
    const int64_t n = (int64_t)I.nx * I.ny * I.nz;
    double s = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * kDiagNT + threadIdx.x; i < n; i += (int64_t)kDiagBlocks * kDiagNT)
        s += current_value(I, field, i);
    const double tot = block_sum<kDiagNT>(s, sred);
    if (threadIdx.x == 0) part[blockIdx.x] = tot;
}

// KH: part[3 b + k] = block sums of vy sin(4 pi x/Nx) w, vy cos(4 pi x/Nx) w and w,
// w = exp(-4 pi |y/Ny - c|), c = 1/4 below y/Ny = 1/2 and 3/4 above; vy from
// the conserved state as getPrimitive computes it (rho = Mass/vol, vy = Momy/rho/vol).
__global__ __launch_bounds__(kDiagNT) void vy_mode_kernel(const Inst* __restrict__ insts, int slot,
                                                          double* __restrict__ part) {
    __shared__ double sred[kDiagNT / 64 + 1];
    CInst& I = ((CInst*)insts)[slot];
    const int64_t n = (int64_t)I.nx * I.ny * I.nz;
    double ss = 0.0, sc = 0.0, sd = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * kDiagNT + threadIdx.x; i < n; i += (int64_t)kDiagBlocks * kDiagNT) {
        const int x = (int)(i % I.nx);
        const int y = (int)((i / I.nx) % I.ny);
        const double rho = current_value(I, 0, i) / I.vol;
        const double vy = current_value(I, 2, i) / rho / I.vol;
        const double xs = (double)x / I.nx, ys = (double)y / I.ny;
        const double w = exp(-4.0 * M_PI * fabs(ys - (ys < 0.5 ? 0.25 : 0.75)));
        ss += vy * sin(4.0 * M_PI * xs) * w;
        sc += vy * cos(4.0 * M_PI * xs) * w;
        sd += w;
    }
    const double a = block_sum<kDiagNT>(ss, sred);
    const double b = block_sum<kDiagNT>(sc, sred);
    const double c = block_sum<kDiagNT>(sd, sred);
    if (threadIdx.x == 0) {
        part[3 * blockIdx.x] = a;
        part[3 * blockIdx.x + 1] = b;
        part[3 * blockIdx.x + 2] = c;
    }
}

// out[v] = sum over b of part[nv b + v], in block order (one thread per value).
__global__ void reduce_parts_kernel(const double* __restrict__ part, int nv, double* __restrict__ out) {
    const int v = threadIdx.x;
    if (v >= nv) return;
    double s = 0.0;
    for (int b = 0; b < kDiagBlocks; b++) s += part[nv * b + v];
    out[v] = s;
}

// Growth front on column x (z = Nz/2): the first y >= Ny/2 with phi < 0.5,
// interpolated with the cell below; 0 when the column never drops below 0.5.
__global__ void growth_front_kernel(const Inst* __restrict__ insts, int slot, const int32_t* __restrict__ xs,
                                    int ncol, double* __restrict__ out) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= ncol) return;
    CInst& I = ((CInst*)insts)[slot];
    const int ix = xs[k];
    const int z = I.nz / 2;
    const int64_t plane = (int64_t)z * I.nx * I.ny;
    double r = 0.0;
    for (int y = I.ny / 2; y < I.ny; y++) {
        const double pv = current_value(I, 0, ix + (int64_t)y * I.nx + plane);
        if (pv < 0.5) {
            const double pm = current_value(I, 0, ix + (int64_t)(y - 1) * I.nx + plane);
            r = (pm - 0.5) / (pm - pv) + y - 1 - I.ny / 2;
            break;
        }
    }
    out[k] = r;
}

}  // namespace dba
