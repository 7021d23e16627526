// coarsening_analysis - droplet count, size distribution and mean radius of
// the RECTILINEAR_GRID outputs of the CH program, 2D or 3D.
//
// Rule of example_Reactive_ModelB/coarsening_analysis.cpp (which is fixed to
// 256^3 and the Z block): threshold avg = (Max + Min) / 2 with Max and Min
// seeded at 0.5 (:118-140); a droplet is a 6-connected (4 in 2D) periodic
// component of cells >= avg started from a cell > avg (:48-111); the volume is
// the number of cells > avg (:145-156); R = cbrt(3 V / (4 pi N)) in 3D
// (:170), and R = sqrt(V / (pi N)) in 2D.
// Differences: dimensions and the scalar block come from the file; visited
// cells are marked in a separate mask (the reference overwrites them with 0,
// which only terminates when avg > 0); the search uses an explicit stack.
//
// Usage: coarsening_analysis [-f FIELD] FILE...      (FIELD: default Z, else C)
//        coarsening_analysis [-f FIELD] -s STEP -n COUNT   (output_<i*STEP>.vtk)
// Appends "i,R" to avg_radius.txt and the droplet sizes of each file, one line
// per file, to size_dist.txt (the reference's two files).
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

struct Field {
    int nx = 0, ny = 0, nz = 0;
    std::vector<double> v;  // x fastest
};

// Reads DIMENSIONS and the named SCALARS block (falls back to the only / first block).
static bool read_field(const std::string& path, const std::string& want, Field* f, std::string* err) {
    std::ifstream in(path);
    if (!in) {
        *err = "can not open " + path;
        return false;
    }
    std::string line;
    std::string first_name;
    bool found = false;
    while (std::getline(in, line)) {
        std::istringstream ss(line);
        std::string key;
        ss >> key;
        if (key == "DIMENSIONS") {
            ss >> f->nx >> f->ny >> f->nz;
        } else if (key == "SCALARS") {
            std::string name;
            ss >> name;
            if (first_name.empty()) first_name = name;
            if (name == want || want.empty()) {
                std::getline(in, line);  // LOOKUP_TABLE
                const size_t n = (size_t)f->nx * f->ny * f->nz;
                f->v.resize(n);
                for (size_t i = 0; i < n; i++)
                    if (!(in >> f->v[i])) {
                        *err = "short SCALARS block in " + path;
                        return false;
                    }
                found = true;
                break;
            }
        }
    }
    if (!found) *err = "no SCALARS " + want + " in " + path;
    return found;
}

struct Result {
    double avg = 0.0;
    int64_t volume = 0;
    std::vector<int64_t> sizes;
};

static Result analyse(const Field& f) {
    Result r;
    double mx = 0.5, mn = 0.5;
    for (double x : f.v) {
        if (x > mx) mx = x;
        if (x < mn) mn = x;
    }
    r.avg = (mx + mn) / 2.;
    for (double x : f.v) r.volume += x > r.avg;
    const int nx = f.nx, ny = f.ny, nz = f.nz;
    const int64_t n = (int64_t)nx * ny * nz;
    std::vector<uint8_t> seen((size_t)n, 0);
    std::vector<int64_t> stack;
    // scan order of the reference: its first array index is the file's slowest axis
    for (int64_t s = 0; s < n; s++) {
        if (seen[s] || !(f.v[s] > r.avg)) continue;
        int64_t size = 0;
        stack.push_back(s);
        seen[s] = 1;
        while (!stack.empty()) {
            const int64_t c = stack.back();
            stack.pop_back();
            size++;
            const int x = (int)(c % nx), y = (int)((c / nx) % ny), z = (int)(c / ((int64_t)nx * ny));
            const int nb[6][3] = {{x - 1, y, z}, {x + 1, y, z}, {x, y - 1, z}, {x, y + 1, z}, {x, y, z - 1}, {x, y, z + 1}};
            for (int k = 0; k < (nz > 1 ? 6 : 4); k++) {
                const int xx = (nb[k][0] + nx) % nx, yy = (nb[k][1] + ny) % ny, zz = (nb[k][2] + nz) % nz;
                const int64_t q = xx + (int64_t)nx * (yy + (int64_t)ny * zz);
                if (seen[q] || f.v[q] < r.avg) continue;
                seen[q] = 1;
                stack.push_back(q);
            }
        }
        r.sizes.push_back(size);
    }
    return r;
}

int main(int argc, char** argv) {
    std::string field = "Z";
    long step = -1, count = -1;
    std::vector<std::string> files;
    for (int i = 1; i < argc; i++) {
        std::string a = argv[i];
        if (a == "-f" && i + 1 < argc) field = argv[++i];
        else if (a == "-s" && i + 1 < argc) step = std::atol(argv[++i]);
        else if (a == "-n" && i + 1 < argc) count = std::atol(argv[++i]);
        else files.push_back(a);
    }
    if (files.empty() && step > 0 && count > 0)
        for (long i = 0; i < count; i++) files.push_back("output_" + std::to_string(i * step) + ".vtk");
    if (files.empty()) {
        std::cout << "usage: coarsening_analysis [-f FIELD] FILE... | [-f FIELD] -s STEP -n COUNT" << std::endl;
        return 1;
    }
    std::ofstream orad("avg_radius.txt", std::ios::app), osz("size_dist.txt", std::ios::app);
    for (size_t i = 0; i < files.size(); i++) {
        Field f;
        std::string err;
        if (!read_field(files[i], field, &f, &err)) {
            std::cout << "Error: " << err << std::endl;
            return 2;
        }
        std::cout << "Read " << files[i] << " success" << std::endl;
        Result r = analyse(f);
        const double N = (double)r.sizes.size();
        const double R = f.nz > 1 ? std::cbrt(r.volume * 0.75 / M_PI / N) : std::sqrt(r.volume / M_PI / N);
        for (int64_t s : r.sizes) osz << s << " ";
        osz << std::endl;
        orad << i << "," << R << std::endl;
        std::cout << "avg = " << r.avg << ", volume = " << r.volume << ", droplets = " << r.sizes.size()
                  << ", R = " << R << std::endl;
    }
    std::cout << "Processing finished." << std::endl;
    return 0;
}
