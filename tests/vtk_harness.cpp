// Host-only harness for the reference-tool tests: writes VTK files with the
// drop-in drivers' own writers (dba_amd/host/host_common.h) from raw FP64
// arrays (little-endian, x fastest), so the reference's analysis programs can
// be run on them without a GPU. Test infrastructure; not part of the library.
//   vtk_harness structured PREFIX T NX NY NZ FIELD.bin      -> PREFIX_T.vtk
//   vtk_harness modelb T NX NY NZ X.bin Y.bin Z.bin PHI.bin  -> output_T.vtk
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

#include "../dba_amd/host/host_common.h"

static std::vector<double> load(const std::string& path, size_t n) {
    std::vector<double> a(n);
    std::ifstream f(path, std::ios::binary);
    f.read(reinterpret_cast<char*>(a.data()), (std::streamsize)(n * sizeof(double)));
    if (!f) {
        std::cerr << "short read: " << path << std::endl;
        std::exit(3);
    }
    return a;
}

int main(int argc, char** argv) {
    if (argc < 2) return 1;
    const std::string mode = argv[1];
    if (mode == "structured" && argc == 8) {
        host::Grid g{std::atoi(argv[4]), std::atoi(argv[5]), std::atoi(argv[6])};
        const std::vector<double> a = load(argv[7], g.n());
        host::write_structured(argv[2], std::atoi(argv[3]), g, a.data());
        return 0;
    }
    if (mode == "modelb" && argc == 10) {
        host::Grid g{std::atoi(argv[3]), std::atoi(argv[4]), std::atoi(argv[5])};
        const std::vector<double> x = load(argv[6], g.n()), y = load(argv[7], g.n()), z = load(argv[8], g.n()),
                                  phi = load(argv[9], g.n());
        std::vector<double> b(g.n());
        for (size_t i = 0; i < g.n(); i++) b[i] = 1.0 - x[i] - y[i] - z[i];  // as the CH driver computes B
        host::write_modelb_output(std::atoi(argv[2]), g, x.data(), y.data(), z.data(), b.data(), phi.data());
        return 0;
    }
    std::cerr << "usage: vtk_harness structured PREFIX T NX NY NZ F.bin | modelb T NX NY NZ X Y Z PHI" << std::endl;
    return 1;
}
