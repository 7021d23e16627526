/*
 * oracle.h - CPU restatement of the reference DBA time step (test infrastructure).
 *
 * THIS IS THE ORACLE, NOT THE PRODUCT. Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg may load it, and only as the checker or as the
 * timed CPU baseline. The HIP path (dba_amd/csrc) never calls it.
 *
 * It follows the reference CPU_version loop structure (full arrays, separate
 * derivative arrays, copy of new into old) and the single-GPU DBA semantics of
 * the reference (awake flags, any() detector, neighbour dilation, stale
 * derived arrays of sleeping blocks), in FP64:
 *   dendrite   example_Dendritic_growth/CPU_version/Base.cpp:106-204, 289-305 (regular)
 *              example_Dendritic_growth/kernel.cu:322-634, 946-987 (DBA)
 *   Model B    example_Reactive_ModelB/CPU_version/base_CPU.cpp:201-344, 563-596 (regular)
 *              example_Reactive_ModelB/kernel.cu:469-763, 1091-1132 (DBA)
 *              example_Reactive_ModelB/Conserved/kernel.cu:526-553, 630-702, 1128-1187
 *   KH         example_Kelvin_Helmholtz/CPU_version/KH_CPU.cpp:263-515, 749-796 (regular)
 *              example_Kelvin_Helmholtz/KH_DBA.cu:418-1044, 1436-1500 (DBA)
 *              example_Kelvin_Helmholtz/Conserved/KH_DBA.cu:1366-1392, 1553-1575
 *
 * Per-function pinning (closed forms, invariants, brute force, golden vectors
 * from the compiled reference CPU_version) lives in tests/test_oracle_*.py.
 */
#ifndef DBA_ORACLE_H
#define DBA_ORACLE_H
#include <stdint.h>
#include "../include/dba/dba.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ora_sim ora_sim;

ora_sim* ora_create(const dba_params* p);
void     ora_destroy(ora_sim* s);
int      ora_set_field(ora_sim* s, int field, const double* src);
int      ora_get_field(ora_sim* s, int field, double* dst);
int      ora_step(ora_sim* s, int nsteps);
/* awake flags for the next step (after dilation), one byte per tile */
int      ora_get_awake(ora_sim* s, int set, uint8_t* dst);
int      ora_set_awake(ora_sim* s, int set, const uint8_t* src);
/* detector output of the last step (before dilation) */
int      ora_get_next(ora_sim* s, int set, uint8_t* dst);
int      ora_ntiles(ora_sim* s);
/* per-step awake counts (set 0, set 1) of executed steps: out[2*i+set] */
int      ora_history(ora_sim* s, int32_t* out, int cap);
double   ora_time(ora_sim* s);
double   ora_last_dt(ora_sim* s);

/* Activation rule pieces, standalone (brute-force pins):
   A[t] = OR over the neighbourhood of N; kind 0 = faces (6 in 3D, z wraps onto
   itself when ntz == 1), kind 1 = 3x3 in-plane incl. diagonals + z faces (KH).
   wrap = 1 periodic tile grid, 0 clamped at edges. */
void ora_dilate(const uint8_t* N, uint8_t* A, int ntx, int nty, int ntz, int kind,
                int wrap_x, int wrap_y, int wrap_z);
/* std::copy_if over ascending tile index; returns count */
int  ora_compact(const uint8_t* A, int nt, int32_t* out);

/* Regular-mode CPU baseline: run nsteps with nthreads y-slab workers
   (nthreads == 1 is the serial loop). Returns wall seconds of the step loop. */
double ora_bench(ora_sim* s, int nsteps, int nthreads);

#ifdef __cplusplus
}
#endif
#endif
