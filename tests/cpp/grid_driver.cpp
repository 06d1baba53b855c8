// grid_driver.cpp -- drives the C++ mirror (include/dragon_gpu.hpp) the way a
// reference user drives Grid2D: fill grid(i,j), set the boundary, initialise
// B, advance, read back.  Used by tests/test_cpp_api.py.
//
//   grid_driver abi
//       checks that need no GPU: bad descriptions throw GpuError(DG_ERR_ARG).
//   grid_driver run DIM NX NY NZ DX DY DZ MHD SOLVER LIMITER CT_ENERGY GAMMA BC T_END IN OUT
//       BC: periodic | outflow.  IN holds the cells (ghosts included, 8
//       doubles each, grid order) and, for MHD in 2D/3D, the faces (3 doubles
//       each); OUT gets the same layout after advance(T_END).
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "dragon_gpu.hpp"

using namespace dragon_gpu;

static int abi_checks() {
    int failures = 0;
    auto expect_arg = [&](const char* what, auto&& make) {
        try {
            make();
            std::printf("FAIL %s: no exception\n", what);
            failures++;
        } catch (const GpuError& e) {
            if (e.code != DG_ERR_ARG) {
                std::printf("FAIL %s: code %d\n", what, e.code);
                failures++;
            }
        }
    };
    Config c;
    expect_arg("nx = 0", [&] { GpuGrid1D g(0, 0.1, 1, c); });
    expect_arg("dx = 0", [&] { GpuGrid2D g(4, 4, 0.0, 0.1, 1, c); });
    Config bad = c;
    bad.muscl = 0;  // CTU needs MUSCL (CTU.cpp:26-30)
    expect_arg("ctu without muscl", [&] { GpuGrid2D g(4, 4, 0.1, 0.1, 1, bad); });
    if (face_mask("XY+") != (X | Y_positive) || face_mask("LR") != X || face_mask("UD") != Y) {
        std::printf("FAIL face_mask\n");
        failures++;
    }
    BoundaryList l = Reflective(Z_negative) + Periodic(Y) + Reflective(Z_positive);
    if (l.items.size() != 3 || l.items[1].faces != Y || l.items[1].kind != DG_BC_PERIODIC) {
        std::printf("FAIL BoundaryList composition\n");
        failures++;
    }
    std::printf(failures ? "abi checks failed\n" : "abi checks passed\n");
    return failures ? 1 : 0;
}

int main(int argc, char** argv) {
    if (argc >= 2 && !std::strcmp(argv[1], "abi")) return abi_checks();
    if (argc != 18 || std::strcmp(argv[1], "run")) {
        std::fprintf(stderr, "usage: grid_driver abi | run DIM NX NY NZ DX DY DZ MHD SOLVER LIMITER CT_ENERGY GAMMA BC T_END IN OUT\n");
        return 2;
    }
    const int dim = std::atoi(argv[2]);
    const int n[3] = {std::atoi(argv[3]), std::atoi(argv[4]), std::atoi(argv[5])};
    const double d[3] = {std::atof(argv[6]), std::atof(argv[7]), std::atof(argv[8])};
    Config c;
    c.mhd = std::atoi(argv[9]);
    c.solver = std::atoi(argv[10]);
    c.limiter = std::atoi(argv[11]);
    c.ct_energy = std::atoi(argv[12]);
    c.gamma = std::atof(argv[13]);
    const std::string bc = argv[14];
    const double t_end = std::atof(argv[15]);
    try {
        GpuGrid g(dim, n, d, 1, c);
        const int G = g.getGhosts();
        const int ex = n[0] + 2 * G, ey = dim >= 2 ? n[1] + 2 * G : 1, ez = dim >= 3 ? n[2] + 2 * G : 1;
        const long cells = (long)ex * ey * ez;
        const bool faces = c.mhd && dim >= 2;
        const long nfaces = faces ? (long)(ex + 1) * (dim >= 2 ? ey + 1 : 1) * (dim >= 3 ? ez + 1 : 1) : 0;
        std::vector<double> in(8 * cells + 3 * nfaces);
        FILE* f = std::fopen(argv[16], "rb");
        if (!f || std::fread(in.data(), sizeof(double), in.size(), f) != in.size()) {
            std::fprintf(stderr, "cannot read %s\n", argv[16]);
            return 2;
        }
        std::fclose(f);
        long m = 0;
        for (int i = -G; i < n[0] + G; i++)
            for (int j = dim >= 2 ? -G : 0; j < (dim >= 2 ? n[1] + G : 1); j++)
                for (int k = dim >= 3 ? -G : 0; k < (dim >= 3 ? n[2] + G : 1); k++, m++)
                    std::memcpy((void*)&g.cell(i, j, k), &in[8 * m], 8 * sizeof(double));
        if (faces) {
            long q = 0;
            for (int i = -G; i <= n[0] + G; i++)
                for (int j = -G; j <= n[1] + G; j++)
                    for (int k = dim >= 3 ? -G : 0; k < (dim >= 3 ? n[2] + G + 1 : 1); k++, q++)
                        std::memcpy((void*)&g.face(i, j, k), &in[8 * cells + 3 * q], 3 * sizeof(double));
        }
        g.push();
        if (bc == "periodic") g.boundary = Periodic();
        else g.boundary = Outflow();
        if (faces) g.initialize_B_fields();
        const long steps = g.advance(t_end);
        g.pull();
        std::vector<double> out(in.size());
        m = 0;
        for (int i = -G; i < n[0] + G; i++)
            for (int j = dim >= 2 ? -G : 0; j < (dim >= 2 ? n[1] + G : 1); j++)
                for (int k = dim >= 3 ? -G : 0; k < (dim >= 3 ? n[2] + G : 1); k++, m++)
                    std::memcpy(&out[8 * m], (const void*)&g.cell(i, j, k), 8 * sizeof(double));
        if (faces) {
            long q = 0;
            for (int i = -G; i <= n[0] + G; i++)
                for (int j = -G; j <= n[1] + G; j++)
                    for (int k = dim >= 3 ? -G : 0; k < (dim >= 3 ? n[2] + G + 1 : 1); k++, q++)
                        std::memcpy(&out[8 * cells + 3 * q], (const void*)&g.face(i, j, k), 3 * sizeof(double));
        }
        f = std::fopen(argv[17], "wb");
        if (!f || std::fwrite(out.data(), sizeof(double), out.size(), f) != out.size()) return 2;
        std::fclose(f);
        std::printf("steps %ld ghosts %d\n", steps, G);
    } catch (const std::exception& e) {
        std::printf("error: %s\n", e.what());
        return 1;
    }
    return 0;
}
