// 3D_dendritic_growth - drop-in driver for example_Dendritic_growth.
//
// Reads the reference Input.txt from the working directory
// (ReadInput, example_Dendritic_growth/kernel.cu:67-111), runs the time loop
// through the DBA C ABI and writes the reference's files:
//   RTValues.out    "Time Vel Act_portion_p Act_portion_T Steps_per_second"
//                   every t_freq steps (Calc_RealTime_Values, kernel.cu:797-855)
//   tip_profile.out growth front at Nx/2-10 .. Nx/2+10 (tip_curvature, kernel.cu:776-790)
// The growth front is evaluated on the device (dba_growth_front); whole fields
// are downloaded only when VTK output is requested.
// Optional "key value" lines after the reference groups:
//   model k94|k93   k94 (default): the reference 3D model; k93: 2D Kobayashi 1993
//                   (anisotropy j, latent heat K, theta0, T_eq; zero-flux walls)
//   tile N          activation tile edge (k94: 4 only; k93: 8, 16 or 32)
//   vtk 1           also write p_<t>.vtk and T_<t>.vtk (STRUCTURED_POINTS, the
//                   layout Compare_VTK and Extract_Tip read) every t_freq steps
//   K, j, theta0, T_eq, radius, T0, bc (periodic|noflux), graph_chunk, threads
//   dump 1          also write the raw fields p_<t>.bin, T_<t>.bin (FP64, x fastest)
#include "host_common.h"

using namespace host;

int main() {
    Clock total;
    int Nx, Ny, Nz, t_total, t_freq, mode;
    double dt, dx, dy, dz, delta, tau, eps, alpha, gamma, Delta, metric_p, metric_T;
    InputReader in("Input.txt");
    in.group(Nx, Ny, Nz);
    in.group(t_total, t_freq, dt);
    in.group(dx, dy, dz);
    in.group(delta, tau, eps, alpha, gamma);
    in.group(Delta);
    in.group(mode);
    in.group(metric_p);
    in.group(metric_T);
    auto kv = in.extras();
    std::cout << "Done with input reading.\n\n";
    const std::string model = gets(kv, "model", "k94");
    const bool k93 = model == "k93";
    if (!k93 && model != "k94") {
        std::cout << "unknown model " << model << " (k94 or k93)" << std::endl;
        return 1;
    }
    std::cout << "Nx = " << Nx << ", Ny = " << Ny << ", Nz = " << Nz << "\n"
              << "t_total = " << t_total << ", t_freq = " << t_freq << ", dt = " << dt << "\n"
              << "delta = " << delta << ", tau = " << tau << ", epsilon = " << eps << ", alpha = " << alpha
              << ", gamma = " << gamma << "\nDelta = " << Delta << "\nMode = " << mode
              << " (0 for regular while 1 for wake-up mode)\nmetric_p = " << metric_p << ", metric_T = " << metric_T
              << "\nmodel = " << model << std::endl;
    if (t_freq <= 0) {
        std::cout << "t_freq must be positive" << std::endl;
        return 1;
    }

    dba_params p{};
    p.model = k93 ? DBA_MODEL_DENDRITE_K93 : DBA_MODEL_DENDRITE_K94;
    p.mode = mode;
    p.nx = Nx;
    p.ny = Ny;
    p.nz = Nz;
    int tile = geti(kv, "tile", k93 ? 16 : 4);
    p.tile_x = p.tile_y = tile;
    p.tile_z = (Nz == 1) ? 1 : tile;
    int bc = gets(kv, "bc", k93 ? "noflux" : "periodic") == "noflux" ? DBA_BC_NOFLUX : DBA_BC_PERIODIC;
    p.bc_x = p.bc_y = p.bc_z = bc;
    p.dx = dx;
    p.dy = dy;
    p.dz = dz;
    p.dt = dt;
    p.delta = delta;
    p.tau = tau;
    p.eps_bar = eps;
    p.alpha = alpha;
    p.gamma = gamma;
    p.thresh_a = metric_p;
    p.thresh_b = metric_T;
    p.aniso_j = geti(kv, "j", 4);
    p.latent_k = getd(kv, "K", 1.6);
    p.theta0 = getd(kv, "theta0", 0.0);
    p.t_eq = getd(kv, "T_eq", 1.0);
    p.threads = geti(kv, "threads", 0);
    const int vtk = geti(kv, "vtk", 0);
    const int dump = geti(kv, "dump", 0);
    const int chunk = geti(kv, "graph_chunk", 100);

    Sim sim(p);
    const Grid& g = sim.g;
    std::cout << "Total device memory: " << [&] {
        int64_t b = 0;
        dba_footprint(&p, &b);
        return b / (1024.0 * 1024.0 * 1024.0);
    }() << " GB" << std::endl;

    // Initialize (example_Dendritic_growth/kernel.cu:167-176): p = 1 - 0.2 dist
    // within dist <= 5 of the centre, T = -Delta. The 2D k93 variant seeds a
    // disk of phi = 1 (radius cells) at the centre and T = T0.
    std::vector<double> phi(g.n()), T(g.n());
    const double radius = getd(kv, "radius", 5.0);
    const double T0 = getd(kv, "T0", 0.0);
    for (int z = 0; z < Nz; z++)
        for (int y = 0; y < Ny; y++)
            for (int x = 0; x < Nx; x++) {
                size_t i = (size_t)x + (size_t)y * Nx + (size_t)z * Nx * Ny;
                int ddx = x - Nx / 2, ddy = y - Ny / 2, ddz = z - Nz / 2;
                if (k93) {
                    phi[i] = (double)(ddx * ddx + ddy * ddy) <= radius * radius ? 1.0 : 0.0;
                    T[i] = T0;
                } else {
                    double dist = std::sqrt((double)(ddx * ddx + ddy * ddy + ddz * ddz));
                    phi[i] = dist <= 5 ? (1. - 0.2 * dist) : 0;
                    T[i] = -Delta;
                }
            }
    sim.upload(DBA_F_PHI, phi);
    sim.upload(DBA_F_T, T);
    if (vtk) {
        write_structured("p", 0, g, phi.data());
        write_structured("T", 0, g, T.data());
    }
    if (dump) {
        write_raw("p_0.bin", phi);
        write_raw("T_0.bin", T);
    }

    Clock loop;
    double stepping_s = 0.0;
    double prev_dist = 0.0;
    int t = 0;
    while (t < t_total) {
        int n = std::min(t_freq - (t % t_freq), t_total - t);
        Clock c;
        check(dba_step_graph(sim.ctx, &sim.id, 1, n, std::min(n, chunk)), "dba_step_graph");
        check(dba_sync(sim.ctx), "dba_sync");
        double sec = c.seconds();
        stepping_s += sec;
        t += n;
        if (t % t_freq) continue;
        std::cout << "Time step: " << t << std::endl;
        // the growth front comes from the device; whole fields only for VTK output
        std::vector<int32_t> cols;
        for (int ix = Nx / 2 - 10; ix <= Nx / 2 + 10; ix++) cols.push_back(((ix % Nx) + Nx) % Nx);
        cols.push_back(Nx / 2);
        std::vector<double> front(cols.size());
        check(dba_growth_front(sim.ctx, sim.id, cols.data(), (int32_t)cols.size(), front.data()), "dba_growth_front");
        int cp = 0, cT = 0;
        sim.awake(DBA_SET_MAIN, &cp);
        sim.awake(DBA_SET_T, &cT);
        double pp = (double)cp / sim.nt, pT = (double)cT / sim.nt;
        printf("awake_p: %f, awake_T: %f\n", pp, pT);
        double dist = front.back();
        double vel = (dist - prev_dist) / t_freq;
        prev_dist = dist;
        std::ostringstream line;
        line << t << " " << vel << " " << pp << " " << pT << " " << (double)t_freq / sec;
        append_line("RTValues.out", "Time Vel Act_portion_p Act_portion_T Steps_per_second", line.str());
        std::ostringstream tip;
        for (size_t k = 0; k + 1 < front.size(); k++) tip << front[k] << " ";
        append_line("tip_profile.out", "", tip.str());
        if (vtk || dump) {
            phi = sim.download(DBA_F_PHI);
            T = sim.download(DBA_F_T);
        }
        if (vtk) {
            write_structured("p", t, g, phi.data());
            write_structured("T", t, g, T.data());
        }
        if (dump) {
            write_raw("p_" + std::to_string(t) + ".bin", phi);
            write_raw("T_" + std::to_string(t) + ".bin", T);
        }
    }
    dba_info info;
    check(dba_instance_info(sim.ctx, sim.id, &info), "dba_instance_info");
    printf("The overall running time is: %f sec.\n", total.seconds());
    printf("The loop running time is: %f sec. %3f percent of overall running time.\n", loop.seconds(),
           loop.seconds() / total.seconds() * 100.);
    printf("In detail: DBA steps (GPU, synchronised) - %f sec, %lld active cell-updates.\n", stepping_s,
           (long long)info.cell_updates);
    return 0;
}
