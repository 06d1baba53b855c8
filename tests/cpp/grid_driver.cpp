// grid_driver.cpp -- drives the C++ mirror (include/dragon_gpu.hpp) the way a
// reference user drives Grid2D: fill grid(i,j), set the boundary, initialise
// B, advance, read back.  Used by tests/test_cpp_api.py.
//
//   grid_driver abi
//       checks that need no GPU: bad descriptions throw GpuError(DG_ERR_ARG).
//   grid_driver hostbc
//       a host-evaluated boundary element against the built-in one it restates,
//       under the unsplit and the split step; a host element's exception
//       reaching the caller; advance() in split mode.
//   grid_driver hoard DIR
//       frames through the C++ DRAGONHOARD mirror (dragon_hoard.hpp): a 3D MHD
//       run written at cycle 7 (cell B in double), restarted into a grid of
//       other spacing, both advanced alike and compared bit for bit; DIR gets
//       the frame and state.bin, the state it holds (cells, ghosts included,
//       8 doubles each, grid order).
//   grid_driver run DIM NX NY NZ DX DY DZ MHD SOLVER LIMITER CT_ENERGY GAMMA BC T_END IN OUT
//       BC: periodic | outflow.  IN holds the cells (ghosts included, 8
//       doubles each, grid order) and, for MHD in 2D/3D, the faces (3 doubles
//       each); OUT gets the same layout after advance(T_END).
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "dragon_gpu.hpp"
#include "dragon_hoard.hpp"

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

// Fixed.cpp's cell fill with corner ghosts, written as a host element.
struct HostFixed : HostGhostFill {
    State st;
    HostFixed(const State& s, int faces) : HostGhostFill(faces, true), st(s) {}
    void apply(GpuGrid& grid) override {
        const int G = grid.getGhosts(), nx = grid.getSizeX(), ny = grid.getSizeY();
        for (int j = -G; j < ny + G; j++)
            for (int i = 1; i <= G; i++) {
                if (faces & X_negative) grid.cell(-i, j) = st;
                if (faces & X_positive) grid.cell(nx - 1 + i, j) = st;
            }
    }
};

// A host element in a boundary list gives the bits of the built-in element it
// restates, under the unsplit and the split step; Config.dimension_unsplit = 0
// makes advance() the split advance.
static int hostbc_checks() {
    Config c;
    c.mhd = 0;
    c.limiter = DG_VANLEER;
    const int n = 48;
    auto fill = [&](GpuGrid2D& g) {
        for (int i = 0; i < n; i++)
            for (int j = 0; j < n; j++) {
                State w;
                w.rho = 1.0 + 0.3 * (j > n / 2);
                w.p = 2.5;
                w.vx = j > n / 2 ? 0.4 : -0.4;
                w.vy = 0.01 * ((i * 7 + j * 3) % 11 - 5);
                g.cell(i, j) = w;
            }
        g.push();
    };
    State inflow;
    inflow.rho = 1.5; inflow.p = 2.5; inflow.vx = 0.3;
    GpuGrid2D a(n, n, 1.0 / n, 1.0 / n, 2, c), b(n, n, 1.0 / n, 1.0 / n, 2, c);
    fill(a);
    fill(b);
    a.boundary = Periodic(Y) + Fixed(inflow, X_negative) + Outflow(X_positive);
    b.boundary = BoundaryList(Periodic(Y)) + std::make_shared<HostFixed>(inflow, X_negative) + Outflow(X_positive);
    const long sa = a.advance(0.02), sb = b.advance(0.02);
    a.pull();
    b.pull();
    int failures = 0;
    const int G = a.getGhosts();
    for (int i = -G; i < n + G; i++)
        for (int j = -G; j < n + G; j++)
            if (std::memcmp(&a.cell(i, j), &b.cell(i, j), sizeof(State))) failures++;
    std::printf("host element vs Fixed: %ld / %ld steps, %d cells differ\n", sa, sb, failures);
    if (sa != sb) failures++;
    // the split step applies the list before every directional sweep (Splitting.cpp:219-299): the host
    // element runs there too, through the boundary callback
    GpuGrid2D a2(n, n, 1.0 / n, 1.0 / n, 2, c), b2(n, n, 1.0 / n, 1.0 / n, 2, c);
    fill(a2);
    fill(b2);
    a2.boundary = Fixed(inflow, X_negative) + Periodic(Y) + Outflow(X_positive);
    b2.boundary = BoundaryList(std::make_shared<HostFixed>(inflow, X_negative)) + Periodic(Y) + Outflow(X_positive);
    const long ta = a2.advance_split(0.02), tb = b2.advance_split(0.02);
    a2.pull();
    b2.pull();
    int dsplit_host = 0;
    for (int i = -G; i < n + G; i++)
        for (int j = -G; j < n + G; j++)
            if (std::memcmp(&a2.cell(i, j), &b2.cell(i, j), sizeof(State))) dsplit_host++;
    std::printf("split step, host element vs Fixed: %ld / %ld steps, %d cells differ\n", ta, tb, dsplit_host);
    failures += dsplit_host + (ta != tb);
    // an exception thrown by a host element stops the call and reaches the caller
    struct Broken : HostGhostFill {
        Broken() : HostGhostFill(X_negative, true) {}
        void apply(GpuGrid&) override { throw std::runtime_error("user boundary failed"); }
    };
    b2.boundary = BoundaryList(Periodic()) + std::make_shared<Broken>();
    for (int split = 0; split < 2; split++) {
        try {
            split ? b2.advance_split(1e-3) : b2.advance(1e-3);
            std::printf("FAIL host element exception swallowed\n");
            failures++;
        } catch (const GpuError&) {
            std::printf("FAIL host element exception turned into GpuError\n");
            failures++;
        } catch (const std::runtime_error& e) {
            if (std::string(e.what()) != "user boundary failed") failures++;
        }
    }
    Config cs = c;
    cs.dimension_unsplit = 0;
    GpuGrid2D s1(n, n, 1.0 / n, 1.0 / n, 2, cs), s2(n, n, 1.0 / n, 1.0 / n, 2, c);
    fill(s1);
    fill(s2);
    s1.boundary = Periodic();
    s2.boundary = Periodic();
    const long k1 = s1.advance(0.02), k2 = s2.advance_split(0.02);
    s1.pull();
    s2.pull();
    int dsplit = 0;
    for (int i = -G; i < n + G; i++)
        for (int j = -G; j < n + G; j++)
            if (std::memcmp(&s1.cell(i, j), &s2.cell(i, j), sizeof(State))) dsplit++;
    std::printf("advance() with dimension_unsplit = 0 vs advance_split: %ld / %ld steps, %d cells differ\n", k1, k2, dsplit);
    failures += dsplit + (k1 != k2);
    std::printf(failures ? "hostbc checks failed\n" : "hostbc checks passed\n");
    return failures ? 1 : 0;
}

// A restart from a frame continues the run bit for bit (main.cpp:71-97).
static int hoard_checks(const std::string& dir) {
    Config c;
    c.limiter = DG_VANLEER;
    const int nx = 12, ny = 10, nz = 14;
    const double pi2 = 6.283185307179586;
    auto make = [&](double dx) { return GpuGrid3D(nx, ny, nz, dx, dx, dx, 3, c); };
    GpuGrid3D a = make(1.0 / 12);
    const int G = a.getGhosts();
    for (int i = 0; i < nx; i++)
        for (int j = 0; j < ny; j++)
            for (int k = 0; k < nz; k++) {
                State w;
                w.rho = 1.0 + 0.2 * std::sin(pi2 * i / nx) * std::cos(pi2 * k / nz);
                w.p = 1.0;
                w.vx = 0.3 * std::sin(pi2 * j / ny);
                w.vy = 0.2 * std::cos(pi2 * k / nz);
                w.vz = 0.1 * std::sin(pi2 * i / nx);
                a.cell(i, j, k) = w;
            }
    for (int i = -G; i < nx + 1 + G; i++)  // each component independent of its own coordinate: div B = 0
        for (int j = -G; j < ny + 1 + G; j++)
            for (int k = -G; k < nz + 1 + G; k++) {
                Vec3& b = a.face(i, j, k);
                b.x = 0.4 + 0.1 * std::sin(pi2 * (j + 0.5) / ny);
                b.y = 0.1 * std::cos(pi2 * (k + 0.5) / nz);
                b.z = 0.1 * std::sin(pi2 * (i + 0.5) / nx);
            }
    a.push();
    a.boundary = Periodic();
    a.initialize_B_fields();
    a.advance(0.02);
    hoard::Options o;
    o.output_dir = dir;
    o.redundant = DGH_WRITE_DOUBLE;  // cell B in double: the restart is exact
    hoard::writeToFile(a, 0.02, 7, "frame_" + hoard::cycle_string(7), o);
    a.pull();
    {
        FILE* f = std::fopen((dir + "/state.bin").c_str(), "wb");
        if (!f) return 1;
        const long nc = (long)(nx + 2 * G) * (ny + 2 * G) * (nz + 2 * G);
        std::fwrite(&a.cell(-G, -G, -G), sizeof(double), 8 * nc, f);
        std::fclose(f);
    }
    GpuGrid3D b = make(0.5);  // spacing replaced by the file's
    b.boundary = Periodic();
    double t = -1;
    int cycle = -1;
    hoard::loadFromFile(b, t, cycle, "frame_" + hoard::cycle_string(7), o);
    int failures = 0;
    if (t != 0.02 || cycle != 7 || b.spacing(0) != a.spacing(0) || b.spacing(2) != a.spacing(2)) {
        std::printf("FAIL restart metadata: t %g cycle %d dx %g\n", t, cycle, b.spacing(0));
        failures++;
    }
    const long sa = a.advance(0.02), sb = b.advance(0.02);
    a.pull();
    b.pull();
    int dc = 0, df = 0;
    for (int i = 0; i < nx; i++)
        for (int j = 0; j < ny; j++)
            for (int k = 0; k < nz; k++)
                if (std::memcmp(&a.cell(i, j, k), &b.cell(i, j, k), sizeof(State))) dc++;
    for (int i = 0; i <= nx; i++)
        for (int j = 0; j <= ny; j++)
            for (int k = 0; k <= nz; k++)
                if (std::memcmp(&a.face(i, j, k), &b.face(i, j, k), sizeof(Vec3))) df++;
    std::printf("restart from frame_00007: %ld / %ld steps after it, %d cells and %d faces differ\n", sa, sb, dc, df);
    failures += dc + df + (sa != sb);
    try {
        hoard::loadFromFile(b, t, cycle, "no_such_frame", o);
        std::printf("FAIL missing file accepted\n");
        failures++;
    } catch (const hoard::HoardError&) {
    }
    std::printf(failures ? "hoard checks failed\n" : "hoard checks passed\n");
    return failures ? 1 : 0;
}

int main(int argc, char** argv) {
    if (argc >= 2 && !std::strcmp(argv[1], "abi")) return abi_checks();
    if (argc >= 2 && !std::strcmp(argv[1], "hostbc")) return hostbc_checks();
    if (argc >= 3 && !std::strcmp(argv[1], "hoard")) return hoard_checks(argv[2]);
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
