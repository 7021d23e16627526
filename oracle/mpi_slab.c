/*
 * mpi_slab.c - MPI slab restatement of the oracle's regular-mode CH loop
 * (TEST INFRASTRUCTURE / CPU BASELINE; only bench.py's cpu_baseline leg and
 * tests/ run it).
 *
 * The decomposition of the reference's MPI_version programs
 * (example_Reactive_ModelB/CPU_version/MPI_version/base_CPU.cpp:182-226: slabs
 * of the slowest axis, a periodic MPI_Sendrecv ring of halo planes before each
 * stencil pass), here in y for the 2D reactive Cahn-Hilliard: each rank owns
 * rows [y0, y1) of the grid and keeps them with two halo rows on each side in a
 * local oracle simulation of (y1 - y0 + 4) rows. Per step: exchange the two
 * halo rows of c, mu on the owned rows plus one ring row (the oracle's
 * ch_calc_mu), c' on the owned rows (ch_update), copy back. Every cell sees
 * the operations of the serial oracle in the same order, so the result is
 * bit-identical to it (tests/test_mpi_slab.py).
 *
 * usage: mpirun -np R mpi_slab PARAMS.bin IC.bin STEPS [OUT.bin]
 *   PARAMS.bin  one dba_params struct (the C ABI's layout), model CH, regular
 *   IC.bin      nx * ny doubles, x fastest
 * Rank 0 prints one JSON line: {"ranks", "steps", "seconds", "cell_updates_per_s"}
 * with the time of the step loop between two barriers (steady clock).
 */
#include <mpi.h>
#include <stdio.h>

#include "oracle.c"

static double now_s(void) {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

int main(int argc, char** argv) {
    MPI_Init(&argc, &argv);
    int rank = 0, size = 1;
    MPI_Comm_rank(MPI_COMM_WORLD, &rank);
    MPI_Comm_size(MPI_COMM_WORLD, &size);
    if (argc < 4) {
        if (rank == 0) fprintf(stderr, "usage: mpi_slab PARAMS.bin IC.bin STEPS [OUT.bin]\n");
        MPI_Abort(MPI_COMM_WORLD, 1);
    }
    dba_params p;
    FILE* f = fopen(argv[1], "rb");
    if (!f || fread(&p, sizeof(p), 1, f) != 1) {
        fprintf(stderr, "rank %d: cannot read %s\n", rank, argv[1]);
        MPI_Abort(MPI_COMM_WORLD, 1);
    }
    fclose(f);
    const int steps = atoi(argv[3]);
    if (p.model != DBA_MODEL_CH || p.mode != DBA_MODE_REGULAR || p.nz != 1 || p.bc_y != DBA_BC_PERIODIC) {
        if (rank == 0) fprintf(stderr, "mpi_slab: regular-mode periodic CH only\n");
        MPI_Abort(MPI_COMM_WORLD, 1);
    }
    const int nx = p.nx, ny = p.ny;
    if (ny < 2 * size) {
        if (rank == 0) fprintf(stderr, "mpi_slab: %d rows for %d ranks\n", ny, size);
        MPI_Abort(MPI_COMM_WORLD, 1);
    }
    /* rows [y0, y1) of this rank: the remainder goes to the first ranks */
    const int base = ny / size, extra = ny % size;
    const int y0 = rank * base + (rank < extra ? rank : extra);
    const int nloc = base + (rank < extra ? 1 : 0);
    dba_params q = p;
    q.ny = nloc + 4;
    q.tile_x = q.tile_y = 1;
    ora_sim* s = ora_create(&q);
    const size_t row = (size_t)nx;
    /* the owned rows from the initial state, at local rows 2 .. nloc + 1 */
    f = fopen(argv[2], "rb");
    if (!f || fseek(f, (long)(sizeof(double) * row * (size_t)y0), SEEK_SET) != 0 ||
        fread(s->f[0] + 2 * row, sizeof(double), row * (size_t)nloc, f) != row * (size_t)nloc) {
        fprintf(stderr, "rank %d: cannot read %s\n", rank, argv[2]);
        MPI_Abort(MPI_COMM_WORLD, 1);
    }
    fclose(f);
    const int up = (rank + 1) % size, down = (rank + size - 1) % size;
    double* c = s->f[0];

    MPI_Barrier(MPI_COMM_WORLD);
    const double t0 = now_s();
    for (int k = 0; k < steps; k++) {
        /* halo rows: my top two rows up (they are the upper rank's rows -2, -1),
           my bottom two rows down (the lower rank's rows nloc, nloc + 1) */
        MPI_Sendrecv(c + (size_t)nloc * row, (int)(2 * row), MPI_DOUBLE, up, 0, c, (int)(2 * row), MPI_DOUBLE, down,
                     0, MPI_COMM_WORLD, MPI_STATUS_IGNORE);
        MPI_Sendrecv(c + 2 * row, (int)(2 * row), MPI_DOUBLE, down, 1, c + (size_t)(nloc + 2) * row, (int)(2 * row),
                     MPI_DOUBLE, up, 1, MPI_COMM_WORLD, MPI_STATUS_IGNORE);
        ch_calc_mu(s, 1, nloc + 3);
        ch_update(s, 2, nloc + 2);
        memcpy(c + 2 * row, s->fn[0] + 2 * row, sizeof(double) * row * (size_t)nloc);
    }
    MPI_Barrier(MPI_COMM_WORLD);
    const double t = now_s() - t0;

    if (argc > 4) {
        /* gather the owned rows in rank order into OUT.bin */
        int* counts = (int*)malloc(sizeof(int) * (size_t)size);
        int* displs = (int*)malloc(sizeof(int) * (size_t)size);
        for (int r = 0; r < size; r++) {
            const int nl = base + (r < extra ? 1 : 0);
            counts[r] = nl * nx;
            displs[r] = (r * base + (r < extra ? r : extra)) * nx;
        }
        double* all = rank == 0 ? (double*)malloc(sizeof(double) * row * (size_t)ny) : NULL;
        MPI_Gatherv(c + 2 * row, nloc * nx, MPI_DOUBLE, all, counts, displs, MPI_DOUBLE, 0, MPI_COMM_WORLD);
        if (rank == 0) {
            FILE* o = fopen(argv[4], "wb");
            if (!o || fwrite(all, sizeof(double), row * (size_t)ny, o) != row * (size_t)ny) {
                fprintf(stderr, "cannot write %s\n", argv[4]);
                MPI_Abort(MPI_COMM_WORLD, 1);
            }
            fclose(o);
            free(all);
        }
        free(counts);
        free(displs);
    }
    if (rank == 0)
        printf("{\"ranks\": %d, \"steps\": %d, \"seconds\": %.9g, \"cell_updates_per_s\": %.9g}\n", size, steps, t,
               (double)nx * ny * steps / t);
    ora_destroy(s);
    MPI_Finalize();
    return 0;
}
