// KH_DBA - drop-in driver for example_Kelvin_Helmholtz (and its Conserved/ variant).
//
// Reads the reference Input.txt from the working directory
// (ReadInput, example_Kelvin_Helmholtz/KH_DBA.cu:88-122), runs the time loop
// through the DBA C ABI and writes the reference's files:
//   output_<mode>_<t>.vtk  rho, p and (vx, vy, 0) at t = 0 and every t_freq steps
//                          (write_output_vtk, KH_DBA.cu:1132-1202)
//   Active_<t>.vtk         per-point block flags, DBA modes (Output_wakeup, :1205-1240)
//   Wakeup_portion.out     "t portion steps/s" (Calc_RealTime_Values, :1242-1270)
//   Integral_mass.out      "t sum(Mass)" (Integral, :1273-1288); Mode 2 (the
//                          Conserved/ program) writes Integral.out with
//                          "t Mass Energy Momx Momy" (Conserved/KH_DBA.cu:1303-1324)
//   VY_mode.out            "t M", the vy growth-mode amplitude (Calc_Vy_mode, :1291-1326)
// Mode: 0 regular, 1 DBA, 2 conserved DBA.
// Optional "key value" lines after the reference groups:
//   gamma G        ratio of specific heats (default 5/3, the reference value)
//   tile N         activation tile edge, 8 (default) or 16
//   walls 1        reflective walls in y (default periodic)
//   metric allgrad|density   wake-up detector (default allgrad, the reference)
//   cfl C          CFL time step dt = C dx / max signal speed (default 0: fixed dt)
//   ic reference|sharp       smooth reference profile (default) or sharp shear layers
//   graph_chunk, threads
//   dump 1         also write the conserved state Mass|Momx|Momy|Energy_<t>.bin (FP64)
#include "host_common.h"

using namespace host;

int main() {
    Clock total;
    int Nx, Ny, Nz, t_total, t_freq, mode;
    double dt, dx, dy, dz, meps;
    InputReader in("Input.txt");
    in.group(Nx, Ny, Nz);
    in.group(t_total, t_freq, dt);
    in.group(dx, dy, dz);
    in.group(mode);
    in.group(meps);
    auto kv = in.extras();
    std::cout << "Done with input reading." << std::endl;
    std::cout << "Nx = " << Nx << ", Ny = " << Ny << ", Nz = " << Nz << "\nt_total = " << t_total
              << ", t_freq = " << t_freq << ", dt = " << dt << "\ndx = " << dx << ", dy = " << dy
              << "\nMode = " << mode << "\nMetric_eps = " << meps << std::endl;
    if (t_freq <= 0 || Nz != 1) {
        std::cout << "t_freq must be positive and Nz must be 1" << std::endl;
        return 1;
    }

    dba_params p{};
    p.model = DBA_MODEL_KH;
    p.mode = mode;
    p.nx = Nx;
    p.ny = Ny;
    p.nz = 1;
    p.tile_x = p.tile_y = geti(kv, "tile", 8);
    p.tile_z = 1;
    p.bc_x = DBA_BC_PERIODIC;
    p.bc_y = geti(kv, "walls", 0) ? DBA_BC_WALL : DBA_BC_PERIODIC;
    p.bc_z = DBA_BC_PERIODIC;
    p.dx = dx;
    p.dy = dy;
    p.dz = dz;
    p.dt = dt;
    p.gas_gamma = getd(kv, "gamma", 5.0 / 3.0);
    p.cfl = getd(kv, "cfl", 0.0);
    p.dt_mode = p.cfl > 0 ? DBA_DT_CFL : DBA_DT_FIXED;
    p.metric_kind = gets(kv, "metric", "allgrad") == "density" ? DBA_KH_METRIC_DENSITY : DBA_KH_METRIC_ALLGRAD;
    p.thresh_a = meps;
    p.threads = geti(kv, "threads", 0);
    const bool sharp = gets(kv, "ic", "reference") == "sharp";
    const int chunk = geti(kv, "graph_chunk", 100);
    const int dump = geti(kv, "dump", 0);
    const double gg = p.gas_gamma, vol = dx * dy;

    Sim sim(p);
    const Grid& g = sim.g;
    const size_t n = g.n();

    // Initial primitives: the smooth profile of KH_DBA.cu:1102-1125 (y / Ny, x / Nx
    // on integer indices), or sharp layers rho = 2, vx = 0.5 for |y - 1/2| < 1/4
    // at cell centres with vy = 0.01 sin(4 pi x) (two Gaussians); p = 2.5.
    std::vector<double> rho(n), vx(n), vy(n), pr(n, 2.5);
    for (int y = 0; y < Ny; y++)
        for (int x = 0; x < Nx; x++) {
            size_t i = (size_t)x + (size_t)y * Nx;
            if (sharp) {
                double xs = (x + 0.5) / Nx, ys = (y + 0.5) / Ny;
                bool inner = std::fabs(ys - 0.5) < 0.25;
                rho[i] = inner ? 2.0 : 1.0;
                vx[i] = inner ? 0.5 : -0.5;
                vy[i] = 0.01 * std::sin(4 * M_PI * xs) *
                        (std::exp(-(ys - 0.25) * (ys - 0.25) / 0.005) + std::exp(-(ys - 0.75) * (ys - 0.75) / 0.005));
                continue;
            }
            const double rho1 = 1.0, rho2 = 2.0, rho_m = 0.5 * (rho1 - rho2);
            const double vx1 = 0.5, vx2 = -0.5, vx_m = 0.5 * (vx1 - vx2);
            double ys = 1.0 * y / Ny;
            if (y < Ny / 4) {
                rho[i] = rho1 - rho_m * std::exp((ys - 0.25) / 0.025);
                vx[i] = vx1 - vx_m * std::exp((ys - 0.25) / 0.025);
            } else if (y < Ny / 2) {
                rho[i] = rho2 + rho_m * std::exp((-ys + 0.25) / 0.025);
                vx[i] = vx2 + vx_m * std::exp((-ys + 0.25) / 0.025);
            } else if (y < 3 * Ny / 4) {
                rho[i] = rho2 + rho_m * std::exp((ys - 0.75) / 0.025);
                vx[i] = vx2 + vx_m * std::exp((ys - 0.75) / 0.025);
            } else {
                rho[i] = rho1 - rho_m * std::exp((-ys + 0.75) / 0.025);
                vx[i] = vx1 - vx_m * std::exp((-ys + 0.75) / 0.025);
            }
            vy[i] = 0.01 * std::sin(4.0 * M_PI * x / Nx);
        }
    // getConserved (KH_DBA.cu:363-367)
    std::vector<std::vector<double>> U(4, std::vector<double>(n));
    for (size_t i = 0; i < n; i++) {
        U[0][i] = rho[i] * vol;
        U[1][i] = rho[i] * vx[i] * vol;
        U[2][i] = rho[i] * vy[i] * vol;
        U[3][i] = (pr[i] / (gg - 1.0) + 0.5 * rho[i] * (vx[i] * vx[i] + vy[i] * vy[i])) * vol;
    }
    for (int k = 0; k < 4; k++) sim.upload(k, U[k]);

    auto primitives = [&]() {  // getPrimitive (KH_DBA.cu:401-405)
        for (size_t i = 0; i < n; i++) {
            rho[i] = U[0][i] / vol;
            vx[i] = U[1][i] / rho[i] / vol;
            vy[i] = U[2][i] / rho[i] / vol;
            pr[i] = (U[3][i] / vol - 0.5 * rho[i] * (vx[i] * vx[i] + vy[i] * vy[i])) * (gg - 1.0);
        }
    };
    auto vtk = [&](int t) {
        if (dump) {  // the conserved state, reference array names
            const char* names[4] = {"Mass", "Momx", "Momy", "Energy"};
            for (int k = 0; k < 4; k++) write_raw(std::string(names[k]) + "_" + std::to_string(t) + ".bin", U[k]);
        }
        write_rectilinear("output_" + std::to_string(mode) + "_" + std::to_string(t) + ".vtk", g,
                          {{"rho", rho.data()}, {"p", pr.data()}}, vx.data(), vy.data());
    };
    auto integral = [&](int t) {
        std::ostringstream o;
        if (mode == DBA_MODE_CONSERVED) {
            double s[4] = {0, 0, 0, 0};
            for (size_t i = 0; i < n; i++)
                for (int k = 0; k < 4; k++) s[k] += U[k][i];
            o << t << " " << s[0] << " " << s[3] << " " << s[1] << " " << s[2];
            append_line("Integral.out", "", o.str());
            std::cout << "Integral of Mass = " << s[0] << "\tIntegral of Energy = " << s[3] << std::endl;
        } else {
            double s = 0.0;
            for (size_t i = 0; i < n; i++) s += U[0][i];
            o << t << " " << s;
            append_line("Integral_mass.out", "", o.str());
            std::cout << "Integral of mass = " << s << std::endl;
        }
    };
    auto vy_mode = [&](int t) {  // Calc_Vy_mode (KH_DBA.cu:1291-1326)
        double ss = 0.0, sc = 0.0, sd = 0.0;
        for (int y = 0; y < Ny; y++)
            for (int x = 0; x < Nx; x++) {
                double xs = (double)x / Nx, ys = (double)y / Ny;
                double c0 = ys < 0.5 ? 0.25 : 0.75;
                double w = std::exp(-4.0 * M_PI * std::fabs(ys - c0));
                double v = vy[(size_t)x + (size_t)y * Nx];
                ss += v * std::sin(4.0 * M_PI * xs) * w;
                sc += v * std::cos(4.0 * M_PI * xs) * w;
                sd += w;
            }
        double M = 2.0 * std::sqrt(ss * ss / (sd * sd) + sc * sc / (sd * sd));
        std::ostringstream o;
        o << t << " " << M;
        append_line("VY_mode.out", "", o.str());
        std::cout << "Growth mode amplitude = " << M << std::endl;
    };

    vtk(0);
    std::cout << "Write initial output file finished." << std::endl;
    integral(0);

    Clock loop;
    double stepping_s = 0.0;
    int t = 0;
    while (t < t_total) {
        int nstep = std::min(t_freq - (t % t_freq), t_total - t);
        Clock c;
        check(dba_step_graph(sim.ctx, &sim.id, 1, nstep, std::min(nstep, chunk)), "dba_step_graph");
        check(dba_sync(sim.ctx), "dba_sync");
        double sec = c.seconds();
        stepping_s += sec;
        t += nstep;
        if (t % t_freq) continue;
        std::cout << "Timestep " << t << std::endl;
        for (int k = 0; k < 4; k++) U[k] = sim.download(k);
        primitives();
        vtk(t);
        int cnt = 0;
        auto mask = sim.awake(DBA_SET_MAIN, &cnt);
        if (mode != DBA_MODE_REGULAR) {
            int nbx = (Nx + p.tile_x - 1) / p.tile_x, nby = (Ny + p.tile_y - 1) / p.tile_y;
            write_wakeup("Active", t, g, p.tile_x, p.tile_y, 1, nbx, nby, {{"wakeup", mask.data()}});
        }
        double portion = (double)cnt / sim.nt;
        std::cout << "Wakeup Portion is " << portion << std::endl;
        std::ostringstream line;
        line << t << " " << portion << " " << (double)t_freq / sec;
        append_line("Wakeup_portion.out", "", line.str());
        integral(t);
        vy_mode(t);
    }
    dba_info info;
    check(dba_instance_info(sim.ctx, sim.id, &info), "dba_instance_info");
    printf("\nThe overall running time is: %f sec.\n", total.seconds());
    printf("The loop running time is: %f sec. %3f percent of overall running time.\n", loop.seconds(),
           loop.seconds() / total.seconds() * 100.);
    printf("DBA steps (GPU, synchronised) - %f sec, %lld active cell-updates, simulated time %g.\n", stepping_s,
           (long long)info.cell_updates, info.sim_time);
    std::cout << std::endl << "Program finished." << std::endl;
    return 0;
}
