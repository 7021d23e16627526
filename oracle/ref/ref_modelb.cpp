// Harness around the reference Model B CPU program (built only where the
// reference tree is mounted). Drives Calc_mu -> Update -> Swap_regular of
// example_Reactive_ModelB/CPU_version/base_CPU.cpp:563-596 on caller arrays.
#include <vector>
#define main dba_reference_modelb_main
#include REF_SOURCE
#undef main

extern "C" int ref_modelb_run(int nx, int ny, int nz, const double* prm, int nsteps, double* x, double* y,
                              double* z, const double* phi) {
    // prm: dx dt chi_xy chi_xz chi_yz chi_xb chi_yb chi_zb vx vy vz rx ry rz n m
    //      epsx epsy epsz epsphi Mx My Mz Mphi K z_crit p_gel k0
    InputPara P;
    P.Nx = nx;
    P.Ny = ny;
    P.Nz = nz;
    P.dx = P.dy = P.dz = prm[0];
    P.dt = prm[1];
    P.chi_xy = prm[2]; P.chi_xz = prm[3]; P.chi_yz = prm[4];
    P.chi_xb = prm[5]; P.chi_yb = prm[6]; P.chi_zb = prm[7];
    P.vx = (int)prm[8]; P.vy = (int)prm[9]; P.vz = (int)prm[10];
    P.rx = (int)prm[11]; P.ry = (int)prm[12]; P.rz = (int)prm[13];
    P.n = (int)prm[14]; P.m = (int)prm[15];
    P.epsilonx_sq = prm[16]; P.epsilony_sq = prm[17]; P.epsilonz_sq = prm[18]; P.epsilonphi_sq = prm[19];
    P.Mobility_x = prm[20]; P.Mobility_y = prm[21]; P.Mobility_z = prm[22]; P.Mobility_phi = prm[23];
    P.K = prm[24]; P.z_crit = prm[25]; P.p_gel = prm[26]; P.k_0 = prm[27];
    size_t n = (size_t)nx * ny * nz;
    std::vector<std::vector<double>> a(9, std::vector<double>(n, 0.0));
    std::vector<double> ph(phi, phi + n), phn(n, 0.0);
    double *b = a[0].data(), *dfdx = a[1].data(), *dfdy = a[2].data(), *dfdz = a[3].data();
    double *xn = a[4].data(), *yn = a[5].data(), *zn = a[6].data();
    for (int t = 0; t < nsteps; t++) {
        Calc_mu(x, y, z, b, ph.data(), dfdx, dfdy, dfdz, P);
        Update(xn, yn, zn, phn.data(), x, y, z, ph.data(), b, dfdx, dfdy, dfdz, P);
        Swap_regular(x, xn, y, yn, z, zn, P);
    }
    return 0;
}
