// CH - drop-in driver for example_Reactive_ModelB (and its Conserved/ variant).
//
// Reads the reference Input.txt from the working directory
// (ReadInput, example_Reactive_ModelB/kernel.cu:85-140), runs the time loop
// through the DBA C ABI and writes the reference's files:
//   output_<t>.vtk  RECTILINEAR_GRID with X, Y, Z, B, PHI at t = 0 and every
//                   t_freq steps (write_output_vtk, kernel.cu:766-879; the layout
//                   coarsening_analysis reads)
//   Wakeup_<t>.vtk  per-point block flags, DBA modes (Output_wakeup, kernel.cu:882-917)
//   RTvalues.out    "Time Act_portion_p Steps_per_second" (Calc_RealTime_Values,
//                   kernel.cu:920-986); Mode 2 (conserved DBA, the Conserved/
//                   program) writes Wakeup_portion.out without a header instead
//                   (Conserved/kernel.cu:966-1000)
//   Integral.out    "t sum_x sum_y sum_z sum_b" (Integral, kernel.cu:989-1009)
// Mode: 0 regular, 1 DBA, 2 conserved DBA.
// Optional "key value" lines after the reference groups:
//   model modelb|ch  modelb (default): the ternary reactive Model B (3D or Nz = 1);
//                    ch: 2D Cahn-Hilliard with linear reaction, one field c
//   tile N           activation tile edge (modelb: 4; ch: 8, 16 or 32)
//   seed S           splitmix64 seed of the initial noise (default 1234)
//   ch only: mobility, kappa, react_gamma, cbar, amp, stencil (5|9)
//   graph_chunk, threads
//   dump 1           also write the raw fields <name>_<t>.bin (FP64, x fastest)
//   droplets 1       droplet analysis of Z (ch: C) on the device at every output:
//                    appends "k,R" to avg_radius.txt and the droplet sizes to
//                    size_dist.txt, k = t / t_freq (the files and rule of
//                    coarsening_analysis.cpp, run on output_<k t_freq>.vtk)
#include "host_common.h"

using namespace host;

int main() {
    Clock total;
    int Nx, Ny, Nz, t_total, t_freq, mode;
    double dt, dx, dy, dz;
    double chi[6], v[3], r[3], nexp, mexp, es[4], mob[4], x0, y0, z0, Keq, zc, pg, k0, meps;
    InputReader in("Input.txt");
    in.group(Nx, Ny, Nz);
    in.group(t_total, t_freq, dt);
    in.group(dx, dy, dz);
    in.group(chi[0], chi[1], chi[2], chi[3], chi[4], chi[5]);
    in.group(v[0], v[1], v[2]);
    in.group(r[0], r[1], r[2]);
    in.group(nexp, mexp);
    in.group(es[0], es[1], es[2], es[3]);
    in.group(mob[0], mob[1], mob[2], mob[3]);
    in.group(x0, y0, z0);
    in.group(Keq, zc, pg);
    in.group(k0);
    in.group(meps);
    in.group(mode);
    auto kv = in.extras();
    std::cout << "Done with input reading.\n\n";
    const std::string model = gets(kv, "model", "modelb");
    const bool ch = model == "ch";
    if (!ch && model != "modelb") {
        std::cout << "unknown model " << model << " (modelb or ch)" << std::endl;
        return 1;
    }
    std::cout << "Nx = " << Nx << ", Ny = " << Ny << ", Nz = " << Nz << "\nt_total = " << t_total
              << ", t_freq = " << t_freq << ", dt = " << dt << "\nMetric_eps = " << meps << "\nMode = " << mode
              << " (0 regular, 1 wake-up, 2 conserved wake-up)\nmodel = " << model << std::endl;
    if (t_freq <= 0) {
        std::cout << "t_freq must be positive" << std::endl;
        return 1;
    }

    dba_params p{};
    p.model = ch ? DBA_MODEL_CH : DBA_MODEL_MODELB;
    p.mode = mode;
    p.nx = Nx;
    p.ny = Ny;
    p.nz = Nz;
    int tile = geti(kv, "tile", ch ? 32 : 4);
    p.tile_x = p.tile_y = tile;
    p.tile_z = Nz == 1 ? 1 : tile;
    p.bc_x = p.bc_y = p.bc_z = DBA_BC_PERIODIC;
    p.dx = dx;
    p.dy = dy;
    p.dz = dz;
    p.dt = dt;
    p.thresh_a = meps;
    p.threads = geti(kv, "threads", 0);
    p.chi_xy = chi[0];
    p.chi_xz = chi[1];
    p.chi_yz = chi[2];
    p.chi_xb = chi[3];
    p.chi_yb = chi[4];
    p.chi_zb = chi[5];
    p.v_x = v[0];
    p.v_y = v[1];
    p.v_z = v[2];
    p.r_x = r[0];
    p.r_y = r[1];
    p.r_z = r[2];
    p.n_exp = nexp;
    p.m_exp = mexp;
    p.epsx_sq = es[0];
    p.epsy_sq = es[1];
    p.epsz_sq = es[2];
    p.epsphi_sq = es[3];
    p.mob_x = mob[0];
    p.mob_y = mob[1];
    p.mob_z = mob[2];
    p.mob_phi = mob[3];
    p.k_eq = Keq;
    p.z_crit = zc;
    p.p_gel = pg;
    p.k_0 = k0;
    p.mobility = getd(kv, "mobility", 1.0);
    p.kappa = getd(kv, "kappa", 1.0);
    p.react_gamma = getd(kv, "react_gamma", 1e-3);
    p.cbar = getd(kv, "cbar", 0.0);
    p.stencil = geti(kv, "stencil", 9);
    const uint64_t seed = (uint64_t)geti(kv, "seed", 1234);
    const int chunk = geti(kv, "graph_chunk", 100);
    const int dump = geti(kv, "dump", 0);
    const int droplets = geti(kv, "droplets", 0);

    Sim sim(p);
    const Grid& g = sim.g;
    const size_t n = g.n();

    // Initialize (example_Reactive_ModelB/kernel.cu:209-227): x = x0 + 0.05 (u - 1/2),
    // y = y0 + 0.05 (u - 1/2), z = z0 + 0.001 (u - 1/2), b = 1 - x - y - z, phi = 0;
    // uniforms from splitmix64 (seeds S, S+1, S+2) instead of cuRAND.
    // ch: c = cbar + amp (2u - 1).
    std::vector<std::vector<double>> f;
    if (ch) {
        const double amp = getd(kv, "amp", 0.05);
        auto u = uniform01(seed, n);
        std::vector<double> c(n);
        for (size_t i = 0; i < n; i++) c[i] = p.cbar + amp * (2.0 * u[i] - 1.0);
        f.push_back(c);
    } else {
        auto ux = uniform01(seed, n), uy = uniform01(seed + 1, n), uz = uniform01(seed + 2, n);
        std::vector<double> X(n), Y(n), Z(n), PHI(n, 0.0);
        for (size_t i = 0; i < n; i++) {
            X[i] = x0 + 0.05 * (ux[i] - 0.5);
            Y[i] = y0 + 0.05 * (uy[i] - 0.5);
            Z[i] = z0 + 0.001 * (uz[i] - 0.5);
        }
        f = {X, Y, Z, PHI};
    }
    for (size_t k = 0; k < f.size(); k++) sim.upload((int)k, f[k]);

    auto analyse = [&](int t) {  // coarsening_analysis.cpp:118-175 on the device
        const int field = ch ? 0 : 2;
        double avg = 0.0;
        int64_t vol = 0, cnt = 0;
        check(dba_droplets(sim.ctx, sim.id, field, &avg, &vol, &cnt, nullptr, 0), "dba_droplets");
        std::vector<int64_t> sizes((size_t)cnt);
        if (cnt) check(dba_droplets(sim.ctx, sim.id, field, &avg, &vol, &cnt, sizes.data(), cnt), "dba_droplets");
        const double N = (double)cnt;
        const double R = Nz > 1 ? std::cbrt(vol * 0.75 / M_PI / N) : std::sqrt(vol / M_PI / N);
        std::ostringstream sz, rad;
        for (int64_t v : sizes) sz << v << " ";
        rad << t / t_freq << "," << R;
        append_line("size_dist.txt", "", sz.str());
        append_line("avg_radius.txt", "", rad.str());
        std::cout << "avg = " << avg << ", volume = " << vol << ", droplets = " << cnt << ", R = " << R << std::endl;
    };
    auto output = [&](int t) {
        if (droplets) analyse(t);
        if (dump) {
            const char* names[4] = {"X", "Y", "Z", "PHI"};
            for (size_t k = 0; k < f.size(); k++)
                write_raw(std::string(ch ? "C" : names[k]) + "_" + std::to_string(t) + ".bin", f[k]);
        }
        if (ch) {
            double s = 0.0;
            for (size_t i = 0; i < n; i++) s += f[0][i];
            append_line("Integral.out", "", std::to_string(t) + " " + [&] {
                std::ostringstream o;
                o << s;
                return o.str();
            }());
            std::cout << "Integral of c = " << s << std::endl;
            write_rectilinear("output_" + std::to_string(t) + ".vtk", g, {{"C", f[0].data()}});
            return;
        }
        std::vector<double> B(n);
        double sx = 0.0, sy = 0.0, sz = 0.0, sb = 0.0;
        for (size_t i = 0; i < n; i++) {
            B[i] = 1.0 - f[0][i] - f[1][i] - f[2][i];
            sx += f[0][i];
            sy += f[1][i];
            sz += f[2][i];
            sb += B[i];
        }
        std::ostringstream o;
        o << t << " " << sx << " " << sy << " " << sz << " " << sb;
        append_line("Integral.out", "", o.str());
        std::cout << "Integral of buffer = " << sb << std::endl;
        write_modelb_output(t, g, f[0].data(), f[1].data(), f[2].data(), B.data(), f[3].data());
    };
    output(0);

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
        int cnt = 0;
        auto mask = sim.awake(DBA_SET_MAIN, &cnt);
        double portion = (double)cnt / sim.nt;
        std::cout << "Wakeup Portion is " << portion << std::endl;
        std::ostringstream line;
        line << t << " " << portion << " " << (double)t_freq / sec;
        if (mode == DBA_MODE_CONSERVED)
            append_line("Wakeup_portion.out", "", line.str());
        else
            append_line("RTvalues.out", "Time Act_portion_p Steps_per_second", line.str());
        for (size_t k = 0; k < f.size(); k++) f[k] = sim.download((int)k);
        output(t);
        if (mode != DBA_MODE_REGULAR) {
            int nbx = (Nx + p.tile_x - 1) / p.tile_x, nby = (Ny + p.tile_y - 1) / p.tile_y;
            write_wakeup("Wakeup", t, g, p.tile_x, p.tile_y, p.tile_z, nbx, nby, {{"wakeup", mask.data()}});
        }
    }
    dba_info info;
    check(dba_instance_info(sim.ctx, sim.id, &info), "dba_instance_info");
    printf("\nPerformance test for %s mode:\n", mode == 0 ? "Regular" : "DBA");
    printf("The overall running time is: %f sec.\n", total.seconds());
    printf("The loop running time is: %f sec. %3f percent of overall running time.\n", loop.seconds(),
           loop.seconds() / total.seconds() * 100.);
    printf("DBA steps (GPU, synchronised) - %f sec, %lld active cell-updates.\n", stepping_s,
           (long long)info.cell_updates);
    return 0;
}
