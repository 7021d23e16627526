// Harness around the reference KH CPU program (built only where the
// reference tree is mounted; output goes to oracle/_ref). It includes the
// reference source unchanged, renames its main, and drives its own step
// functions on caller arrays (example_Kelvin_Helmholtz/CPU_version/KH_CPU.cpp:749-796).
#include <vector>
#define main dba_reference_kh_main
#include REF_SOURCE
#undef main

extern "C" int ref_kh_run(int nx, int ny, double dx, double dt, int nsteps, double* mass, double* momx,
                          double* momy, double* energy) {
    InputPara P;
    P.Nx = nx;
    P.Ny = ny;
    P.Nz = 1;
    P.dx = dx;
    P.dy = dx;
    P.dz = dx;
    P.dt = dt;
    P.t_total = nsteps;
    P.t_freq = nsteps + 1;
    size_t n = (size_t)nx * ny;
    std::vector<std::vector<double>> a(24, std::vector<double>(n, 0.0));
    double *rho = a[0].data(), *p = a[1].data(), *vx = a[2].data(), *vy = a[3].data();
    double *rdx = a[4].data(), *rdy = a[5].data(), *pdx = a[6].data(), *pdy = a[7].data();
    double *vxdx = a[8].data(), *vxdy = a[9].data(), *vydx = a[10].data(), *vydy = a[11].data();
    double *rp = a[12].data(), *pp = a[13].data(), *vxp = a[14].data(), *vyp = a[15].data();
    double *fmx = a[16].data(), *fex = a[17].data(), *fux = a[18].data(), *fvx = a[19].data();
    double *fmy = a[20].data(), *fey = a[21].data(), *fuy = a[22].data(), *fvy = a[23].data();
    for (int t = 0; t < nsteps; t++) {
        getPrimitive(rho, p, vx, vy, mass, energy, momx, momy, P);
        extrap_in_time(rho, p, vx, vy, rdx, rdy, pdx, pdy, vxdx, vxdy, vydx, vydy, rp, pp, vxp, vyp, P);
        calc_flux(rho, p, vx, vy, rdx, rdy, pdx, pdy, vxdx, vxdy, vydx, vydy, rp, pp, vxp, vyp, fmx, fex, fux, fvx,
                  fmy, fey, fuy, fvy, P);
        apply_flux(mass, energy, momx, momy, fmx, fex, fux, fvx, fmy, fey, fuy, fvy, P);
    }
    return 0;
}
