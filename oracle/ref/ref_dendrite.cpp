// Harness around the reference dendrite CPU program (float32, ghost frame,
// built only where the reference tree is mounted). Drives Finite_Diff_p ->
// Finite_Diff_T -> Evolve_p_and_T -> copy of
// example_Dendritic_growth/CPU_version/Base.cpp:289-305 on caller arrays laid
// out with the program's own one-cell ghost frame ((N+2)^3 floats).
#include <vector>
#define main dba_reference_dendrite_main
#include REF_SOURCE
#undef main

extern "C" int ref_dendrite_init(int nx, int ny, int nz, float delta_T, float* p, float* T) {
    InputPara P;
    P.Nx = nx; P.Ny = ny; P.Nz = nz;
    P.Delta = delta_T;
    Initialize(p, T, P);
    return 0;
}

extern "C" int ref_dendrite_run(int nx, int ny, int nz, const float* prm, int nsteps, float* p, float* T) {
    // prm: dx dt delta tau epsilon alpha gamma
    InputPara P;
    P.Nx = nx; P.Ny = ny; P.Nz = nz;
    P.dx = P.dy = P.dz = prm[0];
    P.dt = prm[1];
    P.delta = prm[2]; P.tau = prm[3]; P.epsilon = prm[4]; P.alpha = prm[5]; P.gamma = prm[6];
    size_t n = (size_t)(nx + 2) * (ny + 2) * (nz + 2);
    std::vector<std::vector<float>> a(7, std::vector<float>(n, 0.0f));
    for (int t = 0; t < nsteps; t++) {
        Finite_Diff_p(p, a[0].data(), a[1].data(), a[2].data(), a[3].data(), P);
        Finite_Diff_T(T, a[4].data(), P);
        Evolve_p_and_T(p, a[5].data(), a[0].data(), a[1].data(), a[2].data(), a[3].data(), T, a[6].data(),
                       a[4].data(), P);
        std::memcpy(p, a[5].data(), n * sizeof(float));
        std::memcpy(T, a[6].data(), n * sizeof(float));
    }
    return 0;
}
