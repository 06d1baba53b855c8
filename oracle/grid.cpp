// grid.cpp -- oracle: grids, boundaries, CFL, CT, CTU and the unsplit step.
// TEST INFRASTRUCTURE ONLY (see oracle.hpp).
//
// One function per reference loop nest, same ranges, same operation order.
// Loops over the outermost index may run on CFG.threads threads; each
// iteration is independent, so results do not depend on the thread count.
#include "oracle.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <exception>
#include <thread>

namespace oracle {

static const double k1_8pi = 0.125 / M_PI;

template <class T>
void Arr<T>::clone(const Arr& r, bool ghosts) {
    if (ghosts) { a = r.a; return; }
    if (dim == 1) { for (int i = 0; i < nx; i++) (*this)(i) = r(i); return; }
    if (dim == 2) { for (int i = 0; i < nx; i++) for (int j = 0; j < ny; j++) (*this)(i, j) = r(i, j); return; }
    for (int i = 0; i < nx; i++)
        for (int j = 0; j < ny; j++)
            for (int k = 0; k < nz; k++) (*this)(i, j, k) = r(i, j, k);
}
template struct Arr<Prim>;
template struct Arr<Cons>;
template struct Arr<V3>;

// Parallel loop over [lo, hi).  The first exception in sequential order
// (lowest chunk) is rethrown, which is the one the reference would throw.
template <class F>
static void pfor(int lo, int hi, F&& fn) {
    int T = CFG.threads;
    if (T <= 1 || hi - lo < 2 * T) {
        for (int i = lo; i < hi; i++) fn(i);
        return;
    }
    std::vector<std::thread> th;
    std::vector<std::exception_ptr> ex(T);
    for (int t = 0; t < T; t++) {
        int a = lo + (int)((long)(hi - lo) * t / T), b = lo + (int)((long)(hi - lo) * (t + 1) / T);
        th.emplace_back([&, t, a, b] {
            try { for (int i = a; i < b; i++) fn(i); } catch (...) { ex[t] = std::current_exception(); }
        });
    }
    for (auto& x : th) x.join();
    for (auto& e : ex)
        if (e) std::rethrow_exception(e);
}

static inline Cons cadd(Cons a, const Cons& b) {
    a.rho += b.rho; a.mom.x += b.mom.x; a.mom.y += b.mom.y; a.mom.z += b.mom.z; a.E += b.E;
    if (CFG.mhd) { a.B.x += b.B.x; a.B.y += b.B.y; a.B.z += b.B.z; }
    return a;
}
static inline Cons csub(Cons a, const Cons& b) {
    a.rho -= b.rho; a.mom.x -= b.mom.x; a.mom.y -= b.mom.y; a.mom.z -= b.mom.z; a.E -= b.E;
    if (CFG.mhd) { a.B.x -= b.B.x; a.B.y -= b.B.y; a.B.z -= b.B.z; }
    return a;
}
static inline Cons cmul(Cons a, double s) {
    a.rho *= s; a.mom.x *= s; a.mom.y *= s; a.mom.z *= s; a.E *= s;
    if (CFG.mhd) { a.B.x *= s; a.B.y *= s; a.B.z *= s; }
    return a;
}
static inline V3 cross(const V3& u, const V3& v) {  // Arithmetic.cpp:359-365
    V3 w;
    w.x = u.y * v.z - u.z * v.y;
    w.y = u.z * v.x - u.x * v.z;
    w.z = u.x * v.y - u.y * v.x;
    return w;
}
static inline V3 vneg(V3 a) { a.x *= -1; a.y *= -1; a.z *= -1; return a; }

static std::string fmt2(const char* pre, int i, int j) {
    char b[128]; snprintf(b, sizeof b, "%s(%d,%d)", pre, i, j); return b;
}
static std::string fmt3(const char* pre, int i, int j, int k) {
    char b[128]; snprintf(b, sizeof b, "%s(%d,%d,%d)", pre, i, j, k); return b;
}

// ---------------------------------------------------------------- grid
Grid::Grid(int d, int x, int y, int z, double dx_, double dy_, double dz_, int ghosts, const OConfig& c)
    : dim(d), nx(x), ny(y), nz(z), dx(dx_), dy(dy_), dz(dz_), cfg(c) {
    CFG = cfg;
    g = valid_ghosts(ghosts);
    if (dim == 1) w = Arr<Prim>(nx, g);
    else if (dim == 2) w = Arr<Prim>(nx, ny, g);
    else w = Arr<Prim>(nx, ny, nz, g);
    if (cfg.mhd && dim == 2) B = Arr<V3>(nx + 1, ny + 1, g);
    if (cfg.mhd && dim == 3) B = Arr<V3>(nx + 1, ny + 1, nz + 1, g);
}
int Grid::valid_ghosts(int gg) const {  // Splitting.cpp:19-27
    if (cfg.mhd && cfg.ctu) return std::max(gg, 3);
    if (cfg.muscl) return std::max(gg, 2);
    return std::max(gg, 1);
}

// ---------------------------------------------------------- boundaries
int face_mask(const std::string& s0) {  // GhostFill.cpp:14-49
    int mask = 0;
    std::string s = s0 + " ";
    for (size_t i = 0; i + 1 < s.size(); i++) {
        char c = s[i], n = s[i + 1];
        if (c == 'X' || c == 'x') mask |= (n == '+') ? XP : (n == '-') ? XN : FX;
        else if (c == 'Y' || c == 'y') mask |= (n == '+') ? YP : (n == '-') ? YN : FY;
        else if (c == 'Z' || c == 'z') mask |= (n == '+') ? ZP : (n == '-') ? ZN : FZ;
        else if (c == 'L' || c == 'l') mask |= XN;
        else if (c == 'R' || c == 'r') mask |= XP;
        else if (c == 'U' || c == 'u') mask |= YP;
        else if (c == 'D' || c == 'd') mask |= YN;
    }
    return mask;
}
BC make_bc(int kind, int faces, bool corners, bool gated, const Prim* st) {
    BC b;
    b.kind = kind;
    if (kind == BC_PERIODIC)  // Periodic.cpp:14-15 widens to whole dimensions
        faces = ((faces & FX) ? FX : 0) | ((faces & FY) ? FY : 0) | ((faces & FZ) ? FZ : 0);
    b.faces = faces; b.corners = corners; b.gated = gated;
    if (st) b.state = *st;
    return b;
}

static void outflow(Grid& G, const BC& b) {  // Outflow.cpp
    auto& w = G.w; auto& B = G.B;
    const int ng = G.g, nx = G.nx, ny = G.ny, nz = G.nz, f = b.faces;
    const bool gated = b.gated, mhd = G.cfg.mhd;
    if (G.dim == 1) {
        if (f & XN) for (int g = 1; g <= ng; g++) { w(-g) = w(0); if (gated && w(-g).v.x > 0) w(-g).v.x = 0; }
        if (f & XP) for (int g = 1; g <= ng; g++) { w(nx - 1 + g) = w(nx - 1); if (gated && w(nx - 1 + g).v.x < 0) w(nx - 1 + g).v.x = 0; }
        return;
    }
    int i0 = b.corners ? -ng : 0, in = b.corners ? nx + ng : nx;
    int j0 = i0, jn = b.corners ? ny + ng : ny;
    int k0 = i0, kn = b.corners ? nz + ng : nz;
    if (G.dim == 2) {
        if (f & XN) for (int j = j0; j < jn; j++) for (int g = 1; g <= ng; g++) { w(-g, j) = w(0, j); if (gated && w(-g, j).v.x > 0) w(-g, j).v.x = 0; }
        if (f & XP) for (int j = j0; j < jn; j++) for (int g = 1; g <= ng; g++) { w(nx - 1 + g, j) = w(nx - 1, j); if (gated && w(nx - 1 + g, j).v.x < 0) w(nx - 1 + g, j).v.x = 0; }
        if (f & YN) for (int i = i0; i < in; i++) for (int g = 1; g <= ng; g++) { w(i, -g) = w(i, 0); if (gated && w(i, -g).v.y > 0) w(i, -g).v.y = 0; }
        if (f & YP) for (int i = i0; i < in; i++) for (int g = 1; g <= ng; g++) { w(i, ny - 1 + g) = w(i, ny - 1); if (gated && w(i, ny - 1 + g).v.y < 0) w(i, ny - 1 + g).v.y = 0; }
        if (!mhd) return;
        if (f & XN) for (int j = j0; j <= jn; j++) for (int g = 1; g <= ng; g++) { B(-g, j) = B(0, j); B(-g, j).x = 2 * B(-g + 1, j).x - B(-g + 2, j).x; }
        if (f & XP) for (int j = j0; j <= jn; j++) {
            B(nx, j).y = B(nx - 1, j).y; B(nx, j).z = B(nx - 1, j).z;
            for (int g = 1; g <= ng; g++) { B(nx + g, j) = B(nx, j); B(nx + g, j).x = 2 * B(nx + g - 1, j).x - B(nx + g - 2, j).x; }
        }
        if (f & YN) for (int i = i0; i <= in; i++) for (int g = 1; g <= ng; g++) { B(i, -g) = B(i, 0); B(i, -g).y = 2 * B(i, -g + 1).y - B(i, -g + 2).y; }
        if (f & YP) for (int i = i0; i <= in; i++) {
            B(i, ny).x = B(i, ny - 1).x; B(i, ny).z = B(i, ny - 1).z;
            for (int g = 1; g <= ng; g++) { B(i, ny + g) = B(i, ny); B(i, ny + g).y = 2 * B(i, ny + g - 1).y - B(i, ny + g - 2).y; }
        }
        return;
    }
    if (f & XN) for (int j = j0; j < jn; j++) for (int k = k0; k < kn; k++) for (int g = 1; g <= ng; g++) { w(-g, j, k) = w(0, j, k); if (gated && w(-g, j, k).v.x > 0) w(-g, j, k).v.x = 0; }
    if (f & XP) for (int j = j0; j < jn; j++) for (int k = k0; k < kn; k++) for (int g = 1; g <= ng; g++) { w(nx - 1 + g, j, k) = w(nx - 1, j, k); if (gated && w(nx - 1 + g, j, k).v.x < 0) w(nx - 1 + g, j, k).v.x = 0; }
    if (f & YN) for (int i = i0; i < in; i++) for (int k = k0; k < kn; k++) for (int g = 1; g <= ng; g++) { w(i, -g, k) = w(i, 0, k); if (gated && w(i, -g, k).v.y > 0) w(i, -g, k).v.y = 0; }
    if (f & YP) for (int i = i0; i < in; i++) for (int k = k0; k < kn; k++) for (int g = 1; g <= ng; g++) { w(i, ny - 1 + g, k) = w(i, ny - 1, k); if (gated && w(i, ny - 1 + g, k).v.y < 0) w(i, ny - 1 + g, k).v.y = 0; }
    if (f & ZN) for (int i = i0; i < in; i++) for (int j = j0; j < jn; j++) for (int g = 1; g <= ng; g++) { w(i, j, -g) = w(i, j, 0); if (gated && w(i, j, -g).v.z > 0) w(i, j, -g).v.z = 0; }
    if (f & ZP) for (int i = i0; i < in; i++) for (int j = j0; j < jn; j++) for (int g = 1; g <= ng; g++) { w(i, j, nz - 1 + g) = w(i, j, nz - 1); if (gated && w(i, j, nz - 1 + g).v.z < 0) w(i, j, nz - 1 + g).v.z = 0; }
    if (!mhd) return;
    if (f & XN) for (int j = j0; j <= jn; j++) for (int k = k0; k <= kn; k++) for (int g = 1; g <= ng; g++) { B(-g, j, k) = B(0, j, k); B(-g, j, k).x = 2 * B(-g + 1, j, k).x - B(-g + 2, j, k).x; }
    if (f & XP) for (int j = j0; j <= jn; j++) for (int k = k0; k <= kn; k++) {
        B(nx, j, k).y = B(nx - 1, j, k).y; B(nx, j, k).z = B(nx - 1, j, k).z;
        for (int g = 1; g <= ng; g++) { B(nx + g, j, k) = B(nx, j, k); B(nx + g, j, k).x = 2 * B(nx + g - 1, j, k).x - B(nx + g - 2, j, k).x; }
    }
    if (f & YN) for (int i = i0; i <= in; i++) for (int k = k0; k <= kn; k++) for (int g = 1; g <= ng; g++) { B(i, -g, k) = B(i, 0, k); B(i, -g, k).y = 2 * B(i, -g + 1, k).y - B(i, -g + 2, k).y; }
    if (f & YP) for (int i = i0; i <= in; i++) for (int k = k0; k <= kn; k++) {
        B(i, ny, k).x = B(i, ny - 1, k).x; B(i, ny, k).z = B(i, ny - 1, k).z;
        for (int g = 1; g <= ng; g++) { B(i, ny + g, k) = B(i, ny, k); B(i, ny + g, k).y = 2 * B(i, ny + g - 1, k).y - B(i, ny + g - 2, k).y; }
    }
    if (f & ZN) for (int i = i0; i <= in; i++) for (int j = j0; j <= jn; j++) for (int g = 1; g <= ng; g++) { B(i, j, -g) = B(i, j, 0); B(i, j, -g).z = 2 * B(i, j, -g + 1).z - B(i, j, -g + 2).z; }
    if (f & ZP) for (int i = i0; i <= in; i++) for (int j = j0; j <= jn; j++) {
        B(i, j, nz).x = B(i, j, nz - 1).x; B(i, j, nz).y = B(i, j, nz - 1).y;
        for (int g = 1; g <= ng; g++) { B(i, j, nz + g) = B(i, j, nz); B(i, j, nz + g).z = 2 * B(i, j, nz + g - 1).z - B(i, j, nz + g - 2).z; }
    }
}

static void periodic(Grid& G, const BC& b) {  // Periodic.cpp
    auto& w = G.w; auto& B = G.B;
    const int ng = G.g, nx = G.nx, ny = G.ny, nz = G.nz, f = b.faces;
    if (G.dim == 1) {
        if ((f & FX) == 0) return;
        for (int g = 1; g <= ng; g++) { w(-g) = w(nx - g); w(nx - 1 + g) = w(g - 1); }
        return;
    }
    int i0 = b.corners ? -ng : 0, in = b.corners ? nx + ng : nx;
    int j0 = i0, jn = b.corners ? ny + ng : ny;
    int k0 = i0, kn = b.corners ? nz + ng : nz;
    if (G.dim == 2) {
        if (f & FX) for (int j = j0; j < jn; j++) for (int g = 1; g <= ng; g++) { w(-g, j) = w(nx - g, j); w(nx - 1 + g, j) = w(g - 1, j); }
        if (f & FY) for (int i = i0; i < in; i++) for (int g = 1; g <= ng; g++) { w(i, -g) = w(i, ny - g); w(i, ny - 1 + g) = w(i, g - 1); }
        if (!G.cfg.mhd) return;
        if (f & FX) for (int j = j0; j <= jn; j++) for (int g = 1; g <= ng; g++) { B(-g, j) = B(nx - g, j); B(nx - 1 + g, j) = B(g - 1, j); }
        if (f & FY) for (int i = i0; i <= in; i++) for (int g = 1; g <= ng; g++) { B(i, -g) = B(i, ny - g); B(i, ny - 1 + g) = B(i, g - 1); }
        return;
    }
    if (f & FX) for (int j = j0; j < jn; j++) for (int k = k0; k < kn; k++) for (int g = 1; g <= ng; g++) { w(-g, j, k) = w(nx - g, j, k); w(nx - 1 + g, j, k) = w(g - 1, j, k); }
    if (f & FY) for (int i = i0; i < in; i++) for (int k = k0; k < kn; k++) for (int g = 1; g <= ng; g++) { w(i, -g, k) = w(i, ny - g, k); w(i, ny - 1 + g, k) = w(i, g - 1, k); }
    if (f & FZ) for (int i = i0; i < in; i++) for (int j = j0; j < jn; j++) for (int g = 1; g <= ng; g++) { w(i, j, -g) = w(i, j, nz - g); w(i, j, nz - 1 + g) = w(i, j, g - 1); }
    if (!G.cfg.mhd) return;
    if (f & FX) for (int j = j0; j <= jn; j++) for (int k = k0; k <= kn; k++) for (int g = 1; g <= ng; g++) { B(-g, j, k) = B(nx - g, j, k); B(nx - 1 + g, j, k) = B(g - 1, j, k); }
    if (f & FY) for (int i = i0; i <= in; i++) for (int k = k0; k <= kn; k++) for (int g = 1; g <= ng; g++) { B(i, -g, k) = B(i, ny - g, k); B(i, ny - 1 + g, k) = B(i, g - 1, k); }
    if (f & FZ) for (int i = i0; i <= in; i++) for (int j = j0; j <= jn; j++) for (int g = 1; g <= ng; g++) { B(i, j, -g) = B(i, j, nz - g); B(i, j, nz - 1 + g) = B(i, j, g - 1); }
}

static void reflective(Grid& G, const BC& b) {  // Reflective.cpp
    auto& w = G.w; auto& B = G.B;
    const int ng = G.g, nx = G.nx, ny = G.ny, nz = G.nz, f = b.faces;
    const bool mhd = G.cfg.mhd;
    auto flipB = [&](Prim& s, int d) {
        if (!mhd) return;
        s.B = vneg(s.B);
        if (d == 0) s.B.x *= -1; else if (d == 1) s.B.y *= -1; else s.B.z *= -1;
    };
    if (G.dim == 1) {
        if (f & XN) for (int g = 1; g <= ng; g++) { w(-g) = w(g - 1); w(-g).v.x *= -1; flipB(w(-g), 0); }
        if (f & XP) for (int g = 1; g <= ng; g++) { w(nx - 1 + g) = w(nx - g); w(nx - 1 + g).v.x *= -1; flipB(w(nx - 1 + g), 0); }
        return;
    }
    int i0 = b.corners ? -ng : 0, in = b.corners ? nx + ng : nx;
    int j0 = i0, jn = b.corners ? ny + ng : ny;
    int k0 = i0, kn = b.corners ? nz + ng : nz;
    if (G.dim == 2) {
        if (f & XN) for (int j = j0; j < jn; j++) for (int g = 1; g <= ng; g++) { w(-g, j) = w(g - 1, j); w(-g, j).v.x *= -1; flipB(w(-g, j), 0); }
        if (f & XP) for (int j = j0; j < jn; j++) for (int g = 1; g <= ng; g++) { w(nx - 1 + g, j) = w(nx - g, j); w(nx - 1 + g, j).v.x *= -1; flipB(w(nx - 1 + g, j), 0); }
        if (f & YN) for (int i = i0; i < in; i++) for (int g = 1; g <= ng; g++) { w(i, -g) = w(i, g - 1); w(i, -g).v.y *= -1; flipB(w(i, -g), 1); }
        if (f & YP) for (int i = i0; i < in; i++) for (int g = 1; g <= ng; g++) { w(i, ny - 1 + g) = w(i, ny - g); w(i, ny - 1 + g).v.y *= -1; flipB(w(i, ny - 1 + g), 1); }
        if (!mhd) return;
        if (f & XN) for (int j = j0; j <= jn; j++) for (int g = 1; g <= ng; g++) { B(-g, j) = vneg(B(g - 1, j)); B(-g, j).x = B(g, j).x; }
        if (f & XP) for (int j = j0; j <= jn; j++) {
            B(nx, j) = vneg(B(nx - 1, j));
            for (int g = 1; g <= ng; g++) { B(nx + g, j) = vneg(B(nx - g - 1, j)); B(nx + g, j).x = B(nx - g, j).x; }
        }
        if (f & YN) for (int i = i0; i <= in; i++) for (int g = 1; g <= ng; g++) { B(i, -g) = vneg(B(i, g - 1)); B(i, -g).y = B(i, g).y; }
        if (f & YP) for (int i = i0; i <= in; i++) {
            B(i, ny) = vneg(B(i, ny - 1));
            for (int g = 1; g <= ng; g++) { B(i, ny + g) = vneg(B(i, ny - g - 1)); B(i, ny + g).y = B(i, ny - g).y; }
        }
        return;
    }
    if (f & XN) for (int j = j0; j < jn; j++) for (int k = k0; k < kn; k++) for (int g = 1; g <= ng; g++) { w(-g, j, k) = w(g - 1, j, k); w(-g, j, k).v.x *= -1; flipB(w(-g, j, k), 0); }
    if (f & XP) for (int j = j0; j < jn; j++) for (int k = k0; k < kn; k++) for (int g = 1; g <= ng; g++) { w(nx - 1 + g, j, k) = w(nx - g, j, k); w(nx - 1 + g, j, k).v.x *= -1; flipB(w(nx - 1 + g, j, k), 0); }
    if (f & YN) for (int i = i0; i < in; i++) for (int k = k0; k < kn; k++) for (int g = 1; g <= ng; g++) { w(i, -g, k) = w(i, g - 1, k); w(i, -g, k).v.y *= -1; flipB(w(i, -g, k), 1); }
    if (f & YP) for (int i = i0; i < in; i++) for (int k = k0; k < kn; k++) for (int g = 1; g <= ng; g++) { w(i, ny - 1 + g, k) = w(i, ny - g, k); w(i, ny - 1 + g, k).v.y *= -1; flipB(w(i, ny - 1 + g, k), 1); }
    if (f & ZN) for (int i = i0; i < in; i++) for (int j = j0; j < jn; j++) for (int g = 1; g <= ng; g++) { w(i, j, -g) = w(i, j, g - 1); w(i, j, -g).v.z *= -1; flipB(w(i, j, -g), 2); }
    if (f & ZP) for (int i = i0; i < in; i++) for (int j = j0; j < jn; j++) for (int g = 1; g <= ng; g++) { w(i, j, nz - 1 + g) = w(i, j, nz - g); w(i, j, nz - 1 + g).v.z *= -1; flipB(w(i, j, nz - 1 + g), 2); }
    if (!mhd) return;
    if (f & XN) for (int j = j0; j <= jn; j++) for (int k = k0; k <= kn; k++) for (int g = 1; g <= ng; g++) { B(-g, j, k) = vneg(B(g - 1, j, k)); B(-g, j, k).x = B(g, j, k).x; }
    if (f & XP) for (int j = j0; j <= jn; j++) for (int k = k0; k <= kn; k++) {
        B(nx, j, k) = vneg(B(nx - 1, j, k));
        for (int g = 1; g <= ng; g++) { B(nx + g, j, k) = vneg(B(nx - g - 1, j, k)); B(nx + g, j, k).x = B(nx - g, j, k).x; }
    }
    if (f & YN) for (int i = i0; i <= in; i++) for (int k = k0; k <= kn; k++) for (int g = 1; g <= ng; g++) { B(i, -g, k) = vneg(B(i, g - 1, k)); B(i, -g, k).y = B(i, g, k).y; }
    if (f & YP) for (int i = i0; i <= in; i++) for (int k = k0; k <= kn; k++) {
        B(i, ny, k) = vneg(B(i, ny - 1, k));
        for (int g = 1; g <= ng; g++) { B(i, ny + g, k) = vneg(B(i, ny - g - 1, k)); B(i, ny + g, k).y = B(i, ny - g, k).y; }
    }
    if (f & ZN) for (int i = i0; i <= in; i++) for (int j = j0; j <= jn; j++) for (int g = 1; g <= ng; g++) { B(i, j, -g) = vneg(B(i, j, g - 1)); B(i, j, -g).z = B(i, j, g).z; }
    if (f & ZP) for (int i = i0; i <= in; i++) for (int j = j0; j <= jn; j++) {
        B(i, j, nz) = vneg(B(i, j, nz - 1));
        for (int g = 1; g <= ng; g++) { B(i, j, nz + g) = vneg(B(i, j, nz - g - 1)); B(i, j, nz + g).z = B(i, j, nz - g).z; }
    }
}

static void fixed(Grid& G, const BC& b) {  // Fixed.cpp
    auto& w = G.w; auto& B = G.B;
    const int ng = G.g, nx = G.nx, ny = G.ny, nz = G.nz, f = b.faces;
    const Prim& st = b.state;
    if (G.dim == 1) {
        if (f & XN) for (int g = 1; g <= ng; g++) w(-g) = st;
        if (f & XP) for (int g = 1; g <= ng; g++) w(nx - 1 + g) = st;
        return;
    }
    int i0 = b.corners ? -ng : 0, in = b.corners ? nx + ng : nx;
    int j0 = i0, jn = b.corners ? ny + ng : ny;
    int k0 = i0, kn = b.corners ? nz + ng : nz;
    if (G.dim == 2) {
        if (f & XN) for (int j = j0; j < jn; j++) for (int g = 1; g <= ng; g++) w(-g, j) = st;
        if (f & XP) for (int j = j0; j < jn; j++) for (int g = 1; g <= ng; g++) w(nx - 1 + g, j) = st;
        if (f & YN) for (int i = i0; i < in; i++) for (int g = 1; g <= ng; g++) w(i, -g) = st;
        if (f & YP) for (int i = i0; i < in; i++) for (int g = 1; g <= ng; g++) w(i, ny - 1 + g) = st;
        if (!G.cfg.mhd) return;
        if (f & XN) for (int j = j0; j <= jn; j++) { B(0, j).x = st.B.x; for (int g = 1; g <= ng; g++) B(-g, j) = st.B; }
        if (f & XP) for (int j = j0; j <= jn; j++) for (int g = 0; g < ng; g++) B(nx + g, j) = st.B;
        if (f & YN) for (int i = i0; i <= in; i++) { B(i, 0).y = st.B.y; for (int g = 1; g <= ng; g++) B(i, -g) = st.B; }
        if (f & YP) for (int i = i0; i <= in; i++) for (int g = 0; g < ng; g++) B(i, ny + g) = st.B;
        return;
    }
    if (f & XN) for (int j = j0; j < jn; j++) for (int k = k0; k < kn; k++) for (int g = 1; g <= ng; g++) w(-g, j, k) = st;
    if (f & XP) for (int j = j0; j < jn; j++) for (int k = k0; k < kn; k++) for (int g = 1; g <= ng; g++) w(nx - 1 + g, j, k) = st;
    if (f & YN) for (int i = i0; i < in; i++) for (int k = k0; k < kn; k++) for (int g = 1; g <= ng; g++) w(i, -g, k) = st;
    if (f & YP) for (int i = i0; i < in; i++) for (int k = k0; k < kn; k++) for (int g = 1; g <= ng; g++) w(i, ny - 1 + g, k) = st;
    if (f & ZN) for (int i = i0; i < in; i++) for (int j = j0; j < jn; j++) for (int g = 1; g <= ng; g++) w(i, j, -g) = st;
    if (f & ZP) for (int i = i0; i < in; i++) for (int j = j0; j < jn; j++) for (int g = 1; g <= ng; g++) w(i, j, nz - 1 + g) = st;
    if (!G.cfg.mhd) return;
    if (f & XN) for (int j = j0; j <= jn; j++) for (int k = k0; k <= kn; k++) { B(0, j, k).x = st.B.x; for (int g = 1; g <= ng; g++) B(-g, j, k) = st.B; }
    if (f & XP) for (int j = j0; j <= jn; j++) for (int k = k0; k <= kn; k++) for (int g = 0; g < ng; g++) B(nx + g, j, k) = st.B;
    if (f & YN) for (int i = i0; i <= in; i++) for (int k = k0; k <= kn; k++) { B(i, 0, k).y = st.B.y; for (int g = 1; g <= ng; g++) B(i, -g, k) = st.B; }
    if (f & YP) for (int i = i0; i <= in; i++) for (int k = k0; k <= kn; k++) for (int g = 0; g < ng; g++) B(i, ny + g, k) = st.B;
    if (f & ZN) for (int i = i0; i <= in; i++) for (int j = j0; j <= jn; j++) { B(i, j, 0).z = st.B.z; for (int g = 1; g <= ng; g++) B(i, j, -g) = st.B; }
    if (f & ZP) for (int i = i0; i <= in; i++) for (int j = j0; j <= jn; j++) for (int g = 0; g < ng; g++) B(i, j, nz + g) = st.B;
}

static void apply_one(Grid& G, const BC& b) {
    switch (b.kind) {
        case BC_OUTFLOW: outflow(G, b); break;
        case BC_PERIODIC: periodic(G, b); break;
        case BC_REFLECTIVE: reflective(G, b); break;
        case BC_FIXED: fixed(G, b); break;
        default: break;  // Ignore
    }
}
void Grid::apply_boundary() {  // BoundaryList.hpp:79-104 (implicit outflow first)
    CFG = cfg;
    int uncovered = FALL;
    for (auto& b : boundary) uncovered &= ~b.faces;
    BC imp = make_bc(BC_OUTFLOW, uncovered, true, false);
    apply_one(*this, imp);
    for (auto& b : boundary) apply_one(*this, b);
}

// ------------------------------------------------------------------ CFL
static double signal(const Prim& W) { return CFG.mhd ? c_fast_max(W) : cs(W); }
static double cfl_speed(const Prim& W, double dx, double dy, double dz) {  // CFL.cpp:53-109
    double a = signal(W);
    double speed = 0;
    if (CFG.cfl_mode == CFL_MAX) {
        if (dx > 1e-14) speed = (std::abs(W.v.x) + a) / dx;
        if (dy > 1e-14) speed = std::max(speed, (std::abs(W.v.y) + a) / dy);
        if (dz > 1e-14) speed = std::max(speed, (std::abs(W.v.z) + a) / dz);
    } else if (CFG.cfl_mode == CFL_ADD) {
        if (dx > 1e-14) speed = (std::abs(W.v.x) + a) / dx;
        if (dy > 1e-14) speed += (std::abs(W.v.y) + a) / dy;
        if (dz > 1e-14) speed += (std::abs(W.v.z) + a) / dz;
    } else {
        double p = CFG.cfl_mode;
        if (dx > 1e-14) speed = std::pow((std::abs(W.v.x) + a) / dx, p);
        if (dy > 1e-14) speed += std::pow((std::abs(W.v.y) + a) / dy, p);
        if (dz > 1e-14) speed += std::pow((std::abs(W.v.z) + a) / dz, p);
        speed = std::pow(speed, 1.0 / p);
    }
    return speed;
}
double Grid::cfl_time() const {  // CFL.cpp:113-162
    CFG = cfg;
    double m = 0;
    int n1 = std::min(1, g);
    if (dim == 1) {
        for (int i = -n1; i < nx + n1; i++) m = std::max(m, std::abs(w(i).v.x) + signal(w(i)));
        if (m <= 0) throw std::runtime_error("CFL timestep failed: max signal speed is zero");
        return cfg.cfl_coeff * dx / m;
    }
    std::vector<double> part(nx + 2 * n1, 0.0);
    pfor(-n1, nx + n1, [&](int i) {
        double mm = 0;
        for (int j = -n1; j < ny + n1; j++) {
            if (dim == 2) mm = std::max(mm, cfl_speed(w(i, j), dx, dy, 0.0));
            else for (int k = -n1; k < nz + n1; k++) mm = std::max(mm, cfl_speed(w(i, j, k), dx, dy, dz));
        }
        part[i + n1] = mm;
    });
    for (double p : part) m = std::max(m, p);
    if (m <= 0) throw std::runtime_error("CFL timestep failed: max signal speed is zero");
    return cfg.cfl_coeff / m;
}

// ------------------------------------------------------------------- CT
double upwind_corr(double m, double fm, double fp, double rm, double rp) {  // Electric.cpp:96-101
    if (m > 1e-18) return 0.25 * (fm - rm);
    if (m < -1e-18) return 0.25 * (fp - rp);
    return 0.125 * (fm + fp - rm - rp);
}
static void compute_electric2D(Arr<V3>& E, const Arr<Cons>& FX, const Arr<Cons>& FY, int gg) {  // Electric.cpp:24-35
    const int nx = E.nx - 1, ny = E.ny - 1;
    pfor(-gg, nx + gg + 1, [&](int i) {
        for (int j = -gg; j <= ny + gg; j++) {
            E(i, j).x = -FY(i, j).B.z;
            E(i, j).y = FX(i, j).B.z;
            E(i, j).z = (FY(i - 1, j).B.x - FX(i, j - 1).B.y + FY(i, j).B.x - FX(i, j).B.y) * 0.25;
        }
    });
}
static void compute_electric3D(Arr<V3>& E, const Arr<Cons>& FX, const Arr<Cons>& FY, const Arr<Cons>& FZ, int gg) {  // :38-47
    const int nx = E.nx - 1, ny = E.ny - 1, nz = E.nz - 1;
    pfor(-gg, nx + gg + 1, [&](int i) {
        for (int j = -gg; j <= ny + gg; j++)
            for (int k = -gg; k <= nz + gg; k++) {
                E(i, j, k).x = (FZ(i, j - 1, k).B.y - FY(i, j, k - 1).B.z + FZ(i, j, k).B.y - FY(i, j, k).B.z) * 0.25;
                E(i, j, k).y = (FX(i, j, k - 1).B.z - FZ(i - 1, j, k).B.x + FX(i, j, k).B.z - FZ(i, j, k).B.x) * 0.25;
                E(i, j, k).z = (FY(i - 1, j, k).B.x - FX(i, j - 1, k).B.y + FY(i, j, k).B.x - FX(i, j, k).B.y) * 0.25;
            }
    });
}
static void upwind_electric2D(Arr<V3>& E, const Arr<Cons>& FX, const Arr<Cons>& FY, const Arr<V3>& R, int gg) {  // :52-71
    const int nx = E.nx - 1, ny = E.ny - 1;
    pfor(-gg, nx + gg + 1, [&](int i) {
        for (int j = -gg; j <= ny + gg; j++) {
            V3& e = E(i, j);
            e.z += upwind_corr(FX(i, j).rho, FY(i - 1, j).B.x, FY(i, j).B.x, R(i - 1, j).z, R(i, j).z);
            e.z += upwind_corr(FX(i, j - 1).rho, FY(i - 1, j).B.x, FY(i, j).B.x, R(i - 1, j - 1).z, R(i, j - 1).z);
            e.z += upwind_corr(FY(i, j).rho, -FX(i, j - 1).B.y, -FX(i, j).B.y, R(i, j - 1).z, R(i, j).z);
            e.z += upwind_corr(FY(i - 1, j).rho, -FX(i, j - 1).B.y, -FX(i, j).B.y, R(i - 1, j - 1).z, R(i - 1, j).z);
            e.x += 0.25 * (-2 * FY(i, j).B.z - R(i, j).x - R(i, j - 1).x);
            e.y += 0.25 * (2 * FX(i, j).B.z - R(i, j).y - R(i - 1, j).y);
        }
    });
}
static void upwind_electric3D(Arr<V3>& E, const Arr<Cons>& FX, const Arr<Cons>& FY, const Arr<Cons>& FZ, const Arr<V3>& R, int gg) {  // :74-100
    const int nx = E.nx - 1, ny = E.ny - 1, nz = E.nz - 1;
    pfor(-gg, nx + gg + 1, [&](int i) {
        for (int j = -gg; j <= ny + gg; j++)
            for (int k = -gg; k <= nz + gg; k++) {
                V3& e = E(i, j, k);
                e.y += upwind_corr(FX(i, j, k).rho, -FZ(i - 1, j, k).B.x, -FZ(i, j, k).B.x, R(i - 1, j, k).y, R(i, j, k).y);
                e.y += upwind_corr(FX(i, j, k - 1).rho, -FZ(i - 1, j, k).B.x, -FZ(i, j, k).B.x, R(i - 1, j, k - 1).y, R(i, j, k - 1).y);
                e.z += upwind_corr(FX(i, j, k).rho, FY(i - 1, j, k).B.x, FY(i, j, k).B.x, R(i - 1, j, k).z, R(i, j, k).z);
                e.z += upwind_corr(FX(i, j - 1, k).rho, FY(i - 1, j, k).B.x, FY(i, j, k).B.x, R(i - 1, j - 1, k).z, R(i, j - 1, k).z);
                e.x += upwind_corr(FY(i, j, k).rho, FZ(i, j - 1, k).B.y, FZ(i, j, k).B.y, R(i, j - 1, k).x, R(i, j, k).x);
                e.x += upwind_corr(FY(i, j, k - 1).rho, FZ(i, j - 1, k).B.y, FZ(i, j, k).B.y, R(i, j - 1, k - 1).x, R(i, j, k - 1).x);
                e.z += upwind_corr(FY(i, j, k).rho, -FX(i, j - 1, k).B.y, -FX(i, j, k).B.y, R(i, j - 1, k).z, R(i, j, k).z);
                e.z += upwind_corr(FY(i - 1, j, k).rho, -FX(i, j - 1, k).B.y, -FX(i, j, k).B.y, R(i - 1, j - 1, k).z, R(i - 1, j, k).z);
                e.x += upwind_corr(FZ(i, j, k).rho, -FY(i, j, k - 1).B.z, -FY(i, j, k).B.z, R(i, j, k - 1).x, R(i, j, k).x);
                e.x += upwind_corr(FZ(i, j - 1, k).rho, -FY(i, j, k - 1).B.z, -FY(i, j, k).B.z, R(i, j - 1, k - 1).x, R(i, j - 1, k).x);
                e.y += upwind_corr(FZ(i, j, k).rho, FX(i, j, k - 1).B.z, FX(i, j, k).B.z, R(i, j, k - 1).y, R(i, j, k).y);
                e.y += upwind_corr(FZ(i - 1, j, k).rho, FX(i, j, k - 1).B.z, FX(i, j, k).B.z, R(i - 1, j, k - 1).y, R(i - 1, j, k).y);
            }
    });
}
static void body_electric2D(const Arr<Prim>& w, Arr<V3>& E, int gg) {  // Electric.cpp:108-115
    pfor(-gg - 1, w.nx + gg + 1, [&](int i) {
        for (int j = -gg - 1; j <= w.ny + gg; j++) E(i, j) = cross(w(i, j).B, w(i, j).v);
    });
}
static void body_electric3D(const Arr<Prim>& w, Arr<V3>& E, int gg) {  // :116-125
    pfor(-gg - 1, w.nx + gg + 1, [&](int i) {
        for (int j = -gg - 1; j <= w.ny + gg; j++)
            for (int k = -gg - 1; k <= w.nz + gg; k++) E(i, j, k) = cross(w(i, j, k).B, w(i, j, k).v);
    });
}
void faraday2D(const Arr<V3>& E, Arr<V3>& B, double dt_dx, double dt_dy, int gg) {  // B_fields.cpp:21-40
    const int nx = B.nx - 1, ny = B.ny - 1;
    pfor(-gg, nx + gg + 1, [&](int i) { for (int j = -gg; j < ny + gg; j++) B(i, j).x -= (E(i, j + 1).z - E(i, j).z) * dt_dy; });
    pfor(-gg, nx + gg, [&](int i) { for (int j = -gg; j <= ny + gg; j++) B(i, j).y -= -(E(i + 1, j).z - E(i, j).z) * dt_dx; });
    pfor(-gg, nx + gg, [&](int i) {
        for (int j = -gg; j < ny + gg; j++) B(i, j).z -= (E(i + 1, j).y - E(i, j).y) * dt_dx - (E(i, j + 1).x - E(i, j).x) * dt_dy;
    });
}
void faraday3D(const Arr<V3>& E, Arr<V3>& B, double dt_dx, double dt_dy, double dt_dz, int gg) {  // B_fields.cpp:41-66
    const int nx = B.nx - 1, ny = B.ny - 1, nz = B.nz - 1;
    pfor(-gg, nx + gg + 1, [&](int i) {
        for (int j = -gg; j < ny + gg; j++)
            for (int k = -gg; k < nz + gg; k++)
                B(i, j, k).x -= (E(i, j + 1, k).z - E(i, j, k).z) * dt_dy - (E(i, j, k + 1).y - E(i, j, k).y) * dt_dz;
    });
    pfor(-gg, nx + gg, [&](int i) {
        for (int j = -gg; j <= ny + gg; j++)
            for (int k = -gg; k < nz + gg; k++)
                B(i, j, k).y -= (E(i, j, k + 1).x - E(i, j, k).x) * dt_dz - (E(i + 1, j, k).z - E(i, j, k).z) * dt_dx;
    });
    pfor(-gg, nx + gg, [&](int i) {
        for (int j = -gg; j < ny + gg; j++)
            for (int k = -gg; k <= nz + gg; k++)
                B(i, j, k).z -= (E(i + 1, j, k).y - E(i, j, k).y) * dt_dx - (E(i, j + 1, k).x - E(i, j, k).x) * dt_dy;
    });
}
bool should_protect_thermal(const Prim& w) {  // B_fields.cpp:91-114 (production branches)
    switch (CFG.ct_energy) {
        case CT_TOTAL_E: return false;
        case CT_THERMAL: return true;
        default: {
            double B2 = (w.B.x * w.B.x + w.B.y * w.B.y + w.B.z * w.B.z);
            if (B2 == 0) return false;
            double beta = 8 * M_PI * w.p / (w.B.x * w.B.x + w.B.y * w.B.y + w.B.z * w.B.z);
            return beta <= CFG.ct_beta;
        }
    }
}
static void body_field(const V3& Bc, Prim& w, bool consv_E) {  // B_fields.cpp:70-89
    if (consv_E) {
        Cons U = to_cons(w);
        U.B = Bc;
        w = to_prim(U);
    } else {
        w.B = Bc;
    }
}
void compute_body_fields2D(const Arr<V3>& B, Arr<Prim>& w) {  // B_fields.cpp:129-136
    pfor(-w.g, w.nx + w.g, [&](int i) {
        for (int j = -w.g; j < w.ny + w.g; j++) {
            V3 Bc{(B(i, j).x + B(i + 1, j).x) * 0.5, (B(i, j).y + B(i, j + 1).y) * 0.5, B(i, j).z};
            body_field(Bc, w(i, j), !should_protect_thermal(w(i, j)));
        }
    });
}
void compute_body_fields3D(const Arr<V3>& B, Arr<Prim>& w) {  // B_fields.cpp:152-161
    pfor(-w.g, w.nx + w.g, [&](int i) {
        for (int j = -w.g; j < w.ny + w.g; j++)
            for (int k = -w.g; k < w.nz + w.g; k++) {
                V3 Bc{(B(i, j, k).x + B(i + 1, j, k).x) * 0.5, (B(i, j, k).y + B(i, j + 1, k).y) * 0.5,
                      (B(i, j, k).z + B(i, j, k + 1).z) * 0.5};
                body_field(Bc, w(i, j, k), !should_protect_thermal(w(i, j, k)));
            }
    });
}
void Grid::initialize_B_fields() {  // B_fields.cpp:117-149
    apply_boundary();
    if (dim == 2) {
        for (int i = -g; i < nx + g; i++)
            for (int j = -g; j < ny + g; j++)
                body_field({(B(i, j).x + B(i + 1, j).x) * 0.5, (B(i, j).y + B(i, j + 1).y) * 0.5, B(i, j).z}, w(i, j), false);
    } else if (dim == 3) {
        for (int i = -g; i < nx + g; i++)
            for (int j = -g; j < ny + g; j++)
                for (int k = -g; k < nz + g; k++)
                    body_field({(B(i, j, k).x + B(i + 1, j, k).x) * 0.5, (B(i, j, k).y + B(i, j + 1, k).y) * 0.5,
                                (B(i, j, k).z + B(i, j, k + 1).z) * 0.5}, w(i, j, k), false);
    }
}

// --------------------------------------------------------- half states
// computeHalfStates_X/Y (Unsplit.cpp:292-342) + copyFaceFields (B_fields.cpp:166-183)
static void half_states2D(Arr<Prim>& L, const Arr<Prim>& W, Arr<Prim>& R, const Arr<V3>& B, double dt_dx, double dt_dy, int axis) {
    const int nx = W.nx, ny = W.ny, g = W.g;
    const bool mhd = CFG.mhd;
    if (axis == 0) {
        pfor(-g + 1, nx + g - 1, [&](int i) {
            for (int j = -g; j < ny + g; j++) {
                V3 dB;
                if (mhd) { dB.x = (B(i + 1, j).x - B(i, j).x) * dt_dx; dB.y = (B(i, j + 1).y - B(i, j).y) * dt_dy; }
                muscl(W(i - 1, j), L(i, j), W(i, j), R(i, j), W(i + 1, j), dt_dx, dB);
            }
        });
        for (int j = -g; j < ny + g; j++) {
            L(-g, j) = W(-g, j); R(-g, j) = W(-g, j);
            L(nx - 1 + g, j) = W(nx - 1 + g, j); R(nx - 1 + g, j) = W(nx - 1 + g, j);
        }
        if (mhd)
            for (int i = -g; i < nx + g; i++)
                for (int j = -g; j < ny + g; j++) { L(i, j).B.x = B(i, j).x; R(i, j).B.x = B(i + 1, j).x; }
    } else {
        pfor(-g, nx + g, [&](int i) {
            for (int j = -g + 1; j < ny + g - 1; j++) {
                V3 dB;
                if (mhd) { dB.y = (B(i + 1, j).x - B(i, j).x) * dt_dx; dB.x = (B(i, j + 1).y - B(i, j).y) * dt_dy; }
                Prim l, r;
                muscl(swapXY(W(i, j - 1)), l, swapXY(W(i, j)), r, swapXY(W(i, j + 1)), dt_dy, dB);
                L(i, j) = swapXY(l); R(i, j) = swapXY(r);
            }
        });
        for (int i = -g; i < nx + g; i++) {
            L(i, -g) = W(i, -g); R(i, -g) = W(i, -g);
            L(i, ny - 1 + g) = W(i, ny - 1 + g); R(i, ny - 1 + g) = W(i, ny - 1 + g);
        }
        if (mhd)
            for (int i = -g; i < nx + g; i++)
                for (int j = -g; j < ny + g; j++) { L(i, j).B.y = B(i, j).y; R(i, j).B.y = B(i, j + 1).y; }
    }
}
// Unsplit.cpp:343-436 + B_fields.cpp:184-216
static void half_states3D(Arr<Prim>& L, const Arr<Prim>& W, Arr<Prim>& R, const Arr<V3>& B, double dt_dx, double dt_dy, double dt_dz, int axis) {
    const int nx = W.nx, ny = W.ny, nz = W.nz, g = W.g;
    const bool mhd = CFG.mhd;
    const int lo[3] = {axis == 0 ? -g + 1 : -g, axis == 1 ? -g + 1 : -g, axis == 2 ? -g + 1 : -g};
    const int hi[3] = {axis == 0 ? nx + g - 1 : nx + g, axis == 1 ? ny + g - 1 : ny + g, axis == 2 ? nz + g - 1 : nz + g};
    pfor(lo[0], hi[0], [&](int i) {
        for (int j = lo[1]; j < hi[1]; j++)
            for (int k = lo[2]; k < hi[2]; k++) {
                V3 dB;
                double bx = 0, by = 0, bz = 0;
                if (mhd) {
                    bx = (B(i + 1, j, k).x - B(i, j, k).x) * dt_dx;
                    by = (B(i, j + 1, k).y - B(i, j, k).y) * dt_dy;
                    bz = (B(i, j, k + 1).z - B(i, j, k).z) * dt_dz;
                }
                if (axis == 0) {
                    dB = {bx, by, bz};
                    muscl(W(i - 1, j, k), L(i, j, k), W(i, j, k), R(i, j, k), W(i + 1, j, k), dt_dx, dB);
                } else if (axis == 1) {
                    dB = {by, bx, bz};
                    Prim l, r;
                    muscl(swapXY(W(i, j - 1, k)), l, swapXY(W(i, j, k)), r, swapXY(W(i, j + 1, k)), dt_dy, dB);
                    L(i, j, k) = swapXY(l); R(i, j, k) = swapXY(r);
                } else {
                    dB = {bz, by, bx};
                    Prim l, r;
                    muscl(swapXZ(W(i, j, k - 1)), l, swapXZ(W(i, j, k)), r, swapXZ(W(i, j, k + 1)), dt_dz, dB);
                    L(i, j, k) = swapXZ(l); R(i, j, k) = swapXZ(r);
                }
            }
    });
    if (axis == 0) {
        for (int j = -g; j < ny + g; j++)
            for (int k = -g; k < nz + g; k++) {
                L(-g, j, k) = W(-g, j, k); R(-g, j, k) = W(-g, j, k);
                L(nx - 1 + g, j, k) = W(nx - 1 + g, j, k); R(nx - 1 + g, j, k) = W(nx - 1 + g, j, k);
            }
    } else if (axis == 1) {
        for (int i = -g; i < nx + g; i++)
            for (int k = -g; k < nz + g; k++) {
                L(i, -g, k) = W(i, -g, k); R(i, -g, k) = W(i, -g, k);
                L(i, ny - 1 + g, k) = W(i, ny - 1 + g, k); R(i, ny - 1 + g, k) = W(i, ny - 1 + g, k);
            }
    } else {
        for (int i = -g; i < nx + g; i++)
            for (int j = -g; j < ny + g; j++) {
                L(i, j, -g) = W(i, j, -g); R(i, j, -g) = W(i, j, -g);
                L(i, j, nz - 1 + g) = W(i, j, nz - 1 + g); R(i, j, nz - 1 + g) = W(i, j, nz - 1 + g);
            }
    }
    if (!mhd) return;
    pfor(-g, nx + g, [&](int i) {
        for (int j = -g; j < ny + g; j++)
            for (int k = -g; k < nz + g; k++) {
                if (axis == 0) { L(i, j, k).B.x = B(i, j, k).x; R(i, j, k).B.x = B(i + 1, j, k).x; }
                else if (axis == 1) { L(i, j, k).B.y = B(i, j, k).y; R(i, j, k).B.y = B(i, j + 1, k).y; }
                else { L(i, j, k).B.z = B(i, j, k).z; R(i, j, k).B.z = B(i, j, k + 1).z; }
            }
    });
}

// --------------------------------------------------------------- fluxes
// computeFlux_X/Y/Z (Unsplit.cpp:210-254): F[i] = Riemann(R[i-1], L[i]).
static void flux2D(const Arr<Prim>& L, const Arr<Prim>& R, Arr<Cons>& F, int xl, int xr, int yl, int yr, double dt_d, int axis) {
    if (axis == 0)
        pfor(xl, xr + 1, [&](int i) { for (int j = yl; j < yr; j++) F(i, j) = Riemann(R(i - 1, j), L(i, j)).flux_X(dt_d); });
    else
        pfor(xl, xr, [&](int i) { for (int j = yl; j <= yr; j++) F(i, j) = Riemann(R(i, j - 1), L(i, j)).flux_Y(dt_d); });
}
static void flux3D(const Arr<Prim>& L, const Arr<Prim>& R, Arr<Cons>& F, int xl, int xr, int yl, int yr, int zl, int zr, double dt_d, int axis) {
    if (axis == 0)
        pfor(xl, xr + 1, [&](int i) {
            for (int j = yl; j < yr; j++) for (int k = zl; k < zr; k++) F(i, j, k) = Riemann(R(i - 1, j, k), L(i, j, k)).flux_X(dt_d);
        });
    else if (axis == 1)
        pfor(xl, xr, [&](int i) {
            for (int j = yl; j <= yr; j++) for (int k = zl; k < zr; k++) F(i, j, k) = Riemann(R(i, j - 1, k), L(i, j, k)).flux_Y(dt_d);
        });
    else
        pfor(xl, xr, [&](int i) {
            for (int j = yl; j < yr; j++) for (int k = zl; k <= zr; k++) F(i, j, k) = Riemann(R(i, j, k - 1), L(i, j, k)).flux_Z(dt_d);
        });
}
// applyFluxes (Unsplit.cpp:256-288)
static void apply_fluxes2D(const Arr<Prim>& w, Arr<Prim>& _w, const Arr<Cons>& FX, const Arr<Cons>& FY, double dt_dx, double dt_dy, int gg) {
    _w.clone(w);
    pfor(-gg, w.nx + gg, [&](int i) {
        for (int j = -gg; j < w.ny + gg; j++) {
            Cons U = to_cons(w(i, j));
            U = cadd(U, cmul(csub(FX(i, j), FX(i + 1, j)), dt_dx));
            U = cadd(U, cmul(csub(FY(i, j), FY(i, j + 1)), dt_dy));
            _w(i, j) = to_prim(U);
            if (!is_finite(U)) throw std::runtime_error(fmt2("\tNaN state would be produced at ", i, j) + "\n");
        }
    });
}
static void apply_fluxes3D(const Arr<Prim>& w, Arr<Prim>& _w, const Arr<Cons>& FX, const Arr<Cons>& FY, const Arr<Cons>& FZ,
                           double dt_dx, double dt_dy, double dt_dz, int gg) {
    _w.clone(w);
    pfor(-gg, w.nx + gg, [&](int i) {
        for (int j = -gg; j < w.ny + gg; j++)
            for (int k = -gg; k < w.nz + gg; k++) {
                Cons U = to_cons(w(i, j, k));
                U = cadd(U, cmul(csub(FX(i, j, k), FX(i + 1, j, k)), dt_dx));
                U = cadd(U, cmul(csub(FY(i, j, k), FY(i, j + 1, k)), dt_dy));
                U = cadd(U, cmul(csub(FZ(i, j, k), FZ(i, j, k + 1)), dt_dz));
                _w(i, j, k) = to_prim(U);
                if (!is_finite(U)) throw std::runtime_error(fmt3("\tNaN state would be produced at ", i, j, k) + "\n");
            }
    });
}

// ------------------------------------------------------------------ CTU
// ctu_sweep_hydro 2D + correctState (CTU.cpp:206-218, 264-278)
static void correct_state2D(Arr<Prim>& L, Arr<Prim>& R, const Arr<Cons>& F, double dt_dL, int dimf) {
    const int isX = dimf % 2 == 0 ? 1 : 0, isY = dimf % 2 == 1 ? 1 : 0;
    const int xl = -isY, xr = F.nx + isY, yl = -isX, yr = F.ny + isX;
    pfor(xl, xr, [&](int i) {
        for (int j = yl; j < yr; j++) {
            Cons t = cmul(csub(F(i + isX, j + isY), F(i, j)), dt_dL);
            L(i, j) = to_prim(csub(to_cons(L(i, j)), t));
            R(i, j) = to_prim(csub(to_cons(R(i, j)), t));
        }
    });
}
// correctState 3D (CTU.cpp:280-301)
static void correct_state3D(Arr<Prim>& L, Arr<Prim>& R, const Arr<Cons>& F, double dt_dL, int sdim, int fdim) {
    const int isX = fdim % 3 == 0, isY = fdim % 3 == 1, isZ = fdim % 3 == 2;
    const int isX_ = sdim % 3 == 0, isY_ = sdim % 3 == 1, isZ_ = sdim % 3 == 2;
    const int xl = -isX_, xr = F.nx + isX_, yl = -isY_, yr = F.ny + isY_, zl = -isZ_, zr = F.nz + isZ_;
    pfor(xl, xr, [&](int i) {
        for (int j = yl; j < yr; j++)
            for (int k = zl; k < zr; k++) {
                Cons t = cmul(csub(F(i + isX, j + isY, k + isZ), F(i, j, k)), dt_dL);
                L(i, j, k) = to_prim(csub(to_cons(L(i, j, k)), t));
                R(i, j, k) = to_prim(csub(to_cons(R(i, j, k)), t));
            }
    });
}
// computeCTUFlux_X/Y/Z (CTU.cpp:306-370)
static void ctu_flux3D(const Arr<Prim>& L, const Arr<Prim>& R, const Arr<Cons>& T, Arr<Cons>& F, double dt_d, double dt_t, int axis, int dimt) {
    const int isX = dimt % 3 == 0, isY = dimt % 3 == 1, isZ = dimt % 3 == 2;
    if (axis == 0) {
        const int yl = -1 + isY, yr = F.ny + 1 - isY, zl = -1 + isZ, zr = F.nz + 1 - isZ, xl = 0, xr = F.nx;
        pfor(yl, yr, [&](int j) {
            for (int k = zl; k < zr; k++) {
                Prim pR, pL = to_prim(csub(to_cons(R(xl - 1, j, k)), cmul(csub(T(xl - 1, j + isY, k + isZ), T(xl - 1, j, k)), dt_t)));
                for (int i = xl; i <= xr; i++) {
                    Cons t = cmul(csub(T(i, j + isY, k + isZ), T(i, j, k)), dt_t);
                    pR = to_prim(csub(to_cons(L(i, j, k)), t));
                    F(i, j, k) = Riemann(pL, pR).flux_X(dt_d);
                    pL = to_prim(csub(to_cons(R(i, j, k)), t));
                }
            }
        });
    } else if (axis == 1) {
        const int xl = -1 + isX, xr = F.nx + 1 - isX, yl = 0, yr = F.ny, zl = -1 + isZ, zr = F.nz + 1 - isZ;
        pfor(xl, xr, [&](int i) {
            for (int k = zl; k < zr; k++) {
                Prim pR, pL = to_prim(csub(to_cons(R(i, yl - 1, k)), cmul(csub(T(i + isX, yl - 1, k + isZ), T(i, yl - 1, k)), dt_t)));
                for (int j = yl; j <= yr; j++) {
                    Cons t = cmul(csub(T(i + isX, j, k + isZ), T(i, j, k)), dt_t);
                    pR = to_prim(csub(to_cons(L(i, j, k)), t));
                    F(i, j, k) = Riemann(pL, pR).flux_Y(dt_d);
                    pL = to_prim(csub(to_cons(R(i, j, k)), t));
                }
            }
        });
    } else {
        const int xl = -1 + isX, xr = F.nx + 1 - isX, yl = -1 + isY, yr = F.ny + 1 - isY, zl = 0, zr = F.nz;
        pfor(xl, xr, [&](int i) {
            for (int j = yl; j < yr; j++) {
                Prim pR, pL = to_prim(csub(to_cons(R(i, j, zl - 1)), cmul(csub(T(i + isX, j + isY, zl - 1), T(i, j, zl - 1)), dt_t)));
                for (int k = zl; k <= zr; k++) {
                    Cons t = cmul(csub(T(i + isX, j + isY, k), T(i, j, k)), dt_t);
                    pR = to_prim(csub(to_cons(L(i, j, k)), t));
                    F(i, j, k) = Riemann(pL, pR).flux_Z(dt_d);
                    pL = to_prim(csub(to_cons(R(i, j, k)), t));
                }
            }
        });
    }
}
// ctu_sweep_hydro 3D, 12 solves (CTU.cpp:222-261)
static void ctu_hydro3D(Arr<Prim>& xL, Arr<Prim>& xR, Arr<Prim>& yL, Arr<Prim>& yR, Arr<Prim>& zL, Arr<Prim>& zR,
                        double dt_dx, double dt_dy, double dt_dz) {
    const int nx = xL.nx, ny = xL.ny, nz = xL.nz, g = xL.g;
    Arr<Cons> FX(nx, ny, nz, g), FY(nx, ny, nz, g), FZ(nx, ny, nz, g), FXz(nx, ny, nz, g), FYz(nx, ny, nz, g);
    flux3D(xL, xR, FX, 0, nx, -1, ny + 1, -1, nz + 1, dt_dx, 0);
    flux3D(yL, yR, FY, -1, nx + 1, 0, ny, -1, nz + 1, dt_dy, 1);
    flux3D(zL, zR, FZ, -1, nx + 1, -1, ny + 1, 0, nz, dt_dz, 2);
    ctu_flux3D(xL, xR, FZ, FXz, dt_dx, (dt_dz / 3.0), 0, 2);
    ctu_flux3D(yL, yR, FZ, FYz, dt_dy, (dt_dz / 3.0), 1, 2);
    // F_Xy reuses F_Z's storage; each [i,j,k] is read before it is written
    Arr<Cons>& FXy = FZ;
    {
        Arr<Cons> tmp = FZ;  // separate output keeps the read-before-write order exact
        ctu_flux3D(xL, xR, FY, tmp, dt_dx, (dt_dy / 3.0), 0, 1);
        FXy = tmp;
    }
    Arr<Cons>& FZy = FY;
    {
        Arr<Cons> tmp = FY;
        ctu_flux3D(zL, zR, FY, tmp, dt_dz, (dt_dy / 3.0), 2, 1);
        FZy = tmp;
    }
    correct_state3D(xL, xR, FYz, (0.5 * dt_dy), 0, 1);
    correct_state3D(xL, xR, FZy, (0.5 * dt_dz), 0, 2);
    Arr<Cons>& FYx = FYz;
    {
        Arr<Cons> tmp = FYz;
        ctu_flux3D(yL, yR, FX, tmp, dt_dy, (dt_dx / 3.0), 1, 0);
        FYx = tmp;
    }
    Arr<Cons>& FZx = FZy;
    {
        Arr<Cons> tmp = FZy;
        ctu_flux3D(zL, zR, FX, tmp, dt_dz, (dt_dx / 3.0), 2, 0);
        FZx = tmp;
    }
    correct_state3D(yL, yR, FXz, (0.5 * dt_dx), 1, 0);
    correct_state3D(yL, yR, FZx, (0.5 * dt_dz), 1, 2);
    correct_state3D(zL, zR, FXy, (0.5 * dt_dx), 2, 0);
    correct_state3D(zL, zR, FYx, (0.5 * dt_dy), 2, 1);
}

// ctu_sweep_MHD 2D (CTU.cpp:133-203)
static void ctu_mhd2D(Arr<Prim>& xL, Arr<Prim>& xR, Arr<Prim>& yL, Arr<Prim>& yR, const Arr<V3>& B, const Arr<Prim>& w,
                      Arr<V3>& E, double dt_dx, double dt_dy) {
    const int nx = xL.nx, ny = xL.ny, g = xL.g;
    Arr<Cons> FX(nx, ny, g), FY(nx, ny, g);
    flux2D(xL, xR, FX, -1, nx + 1, -2, ny + 2, dt_dx, 0);
    flux2D(yL, yR, FY, -2, nx + 2, -1, ny + 1, dt_dy, 1);
    Arr<V3> _B(nx + 1, ny + 1, g), E0(nx + 1, ny + 1, g);
    compute_electric2D(E0, FX, FY, 1);
    body_electric2D(w, E, 1);
    upwind_electric2D(E0, FX, FY, E, 1);
    _B.clone(B);
    faraday2D(E0, _B, 0.5 * dt_dx, 0.5 * dt_dy, 1);
    {
        Arr<Prim> whalf(nx, ny, g);
        apply_fluxes2D(w, whalf, FX, FY, 0.5 * dt_dx, 0.5 * dt_dy, 1);
        compute_body_fields2D(_B, whalf);
        body_electric2D(whalf, E, 0);
    }
    pfor(-1, nx + 1, [&](int i) {
        for (int j = -1; j < ny + 1; j++) {
            double dBx = (B(i + 1, j).x - B(i, j).x) * dt_dx;
            double dBy = (B(i, j + 1).y - B(i, j).y) * dt_dy;
            const Prim& W = w(i, j);
            Cons corr = cmul(cmul(csub(FY(i, j), FY(i, j + 1)), 0.5), dt_dy);
            corr.B = V3{};
            double lim_z = minmod(-dBy, dBx);
            corr.E += k1_8pi * (W.B.z * W.v.z * lim_z);
            corr.B.z = 0.5 * dt_dy * (E0(i, j + 1).x - E0(i, j).x) + 0.5 * W.v.z * lim_z;
            Cons uL = cadd(to_cons(xL(i, j)), corr), uR = cadd(to_cons(xR(i, j)), corr);
            uL.B.x = _B(i, j).x;
            uR.B.x = _B(i + 1, j).x;
            xL(i, j) = to_prim(uL); xR(i, j) = to_prim(uR);

            corr = cmul(cmul(csub(FX(i, j), FX(i + 1, j)), 0.5), dt_dx);
            corr.B = V3{};
            lim_z = minmod(-dBx, dBy);
            corr.E += k1_8pi * (W.B.z * W.v.z * lim_z);
            corr.B.z = -0.5 * dt_dx * (E0(i + 1, j).y - E0(i, j).y) + 0.5 * W.v.z * lim_z;
            uL = cadd(to_cons(yL(i, j)), corr); uR = cadd(to_cons(yR(i, j)), corr);
            uL.B.y = _B(i, j).y;
            uR.B.y = _B(i, j + 1).y;
            yL(i, j) = to_prim(uL); yR(i, j) = to_prim(uR);
        }
    });
}

// ctu_sweep_MHD 3D (CTU.cpp:34-130)
static void ctu_mhd3D(Arr<Prim>& xL, Arr<Prim>& xR, Arr<Prim>& yL, Arr<Prim>& yR, Arr<Prim>& zL, Arr<Prim>& zR,
                      const Arr<V3>& B, const Arr<Prim>& w, Arr<V3>& E, double dt_dx, double dt_dy, double dt_dz) {
    const int nx = xL.nx, ny = xL.ny, nz = xL.nz, g = xL.g;
    Arr<Cons> FX(nx, ny, nz, g), FY(nx, ny, nz, g), FZ(nx, ny, nz, g);
    flux3D(xL, xR, FX, -1, nx + 1, -2, ny + 2, -2, nz + 2, dt_dx, 0);
    flux3D(yL, yR, FY, -2, nx + 2, -1, ny + 1, -2, nz + 2, dt_dy, 1);
    flux3D(zL, zR, FZ, -2, nx + 2, -2, ny + 2, -1, nz + 1, dt_dz, 2);
    Arr<V3> _B(nx + 1, ny + 1, nz + 1, g), E0(nx + 1, ny + 1, nz + 1, g);
    compute_electric3D(E0, FX, FY, FZ, 1);
    body_electric3D(w, E, 1);
    upwind_electric3D(E0, FX, FY, FZ, E, 1);
    _B.clone(B);
    faraday3D(E0, _B, 0.5 * dt_dx, 0.5 * dt_dy, 0.5 * dt_dz, 1);
    {
        Arr<Prim> whalf(nx, ny, nz, g);
        apply_fluxes3D(w, whalf, FX, FY, FZ, 0.5 * dt_dx, 0.5 * dt_dy, 0.5 * dt_dz, 1);
        compute_body_fields3D(_B, whalf);
        body_electric3D(whalf, E, 0);
    }
    pfor(-1, nx + 1, [&](int i) {
        for (int j = -1; j < ny + 1; j++)
            for (int k = -1; k < nz + 1; k++) {
                double dBx = (B(i + 1, j, k).x - B(i, j, k).x) * dt_dx;
                double dBy = (B(i, j + 1, k).y - B(i, j, k).y) * dt_dy;
                double dBz = (B(i, j, k + 1).z - B(i, j, k).z) * dt_dz;
                const Prim& W = w(i, j, k);
                Cons uL, uR;
                // x half states
                Cons corr = cmul(cadd(cmul(csub(FY(i, j, k), FY(i, j + 1, k)), dt_dy), cmul(csub(FZ(i, j, k), FZ(i, j, k + 1)), dt_dz)), 0.5);
                corr.B = V3{};
                double lim_y = minmod(-dBz, dBx);
                double lim_z = minmod(-dBy, dBx);
                corr.E += k1_8pi * (W.B.y * W.v.y * lim_y + W.B.z * W.v.z * lim_z);
                corr.B.y = -0.25 * dt_dz * (E0(i, j, k + 1).x - E0(i, j, k).x + E0(i, j + 1, k + 1).x - E0(i, j + 1, k).x) + 0.5 * W.v.y * lim_y;
                corr.B.z = 0.25 * dt_dy * (E0(i, j + 1, k).x - E0(i, j, k).x + E0(i, j + 1, k + 1).x - E0(i, j, k + 1).x) + 0.5 * W.v.z * lim_z;
                uL = cadd(to_cons(xL(i, j, k)), corr); uR = cadd(to_cons(xR(i, j, k)), corr);
                uL.B.x = _B(i, j, k).x; uR.B.x = _B(i + 1, j, k).x;
                xL(i, j, k) = to_prim(uL); xR(i, j, k) = to_prim(uR);
                // y half states
                corr = cmul(cadd(cmul(csub(FX(i, j, k), FX(i + 1, j, k)), dt_dx), cmul(csub(FZ(i, j, k), FZ(i, j, k + 1)), dt_dz)), 0.5);
                corr.B = V3{};
                double lim_x = minmod(-dBz, dBy);
                lim_z = minmod(-dBx, dBy);
                corr.E += k1_8pi * (W.B.x * W.v.x * lim_x + W.B.z * W.v.z * lim_z);
                corr.B.x = 0.25 * dt_dz * (E0(i, j, k + 1).y - E0(i, j, k).y + E0(i + 1, j, k + 1).y - E0(i + 1, j, k).y) + 0.5 * W.v.x * lim_x;
                corr.B.z = -0.25 * dt_dx * (E0(i + 1, j, k).y - E0(i, j, k).y + E0(i + 1, j, k + 1).y - E0(i, j, k + 1).y) + 0.5 * W.v.z * lim_z;
                uL = cadd(to_cons(yL(i, j, k)), corr); uR = cadd(to_cons(yR(i, j, k)), corr);
                uL.B.y = _B(i, j, k).y; uR.B.y = _B(i, j + 1, k).y;
                yL(i, j, k) = to_prim(uL); yR(i, j, k) = to_prim(uR);
                // z half states
                corr = cmul(cadd(cmul(csub(FX(i, j, k), FX(i + 1, j, k)), dt_dx), cmul(csub(FY(i, j, k), FY(i, j + 1, k)), dt_dy)), 0.5);
                corr.B = V3{};
                lim_x = minmod(-dBy, dBz);
                lim_y = minmod(-dBx, dBz);
                corr.E += k1_8pi * (W.B.x * W.v.x * lim_x + W.B.y * W.v.y * lim_y);
                corr.B.x = -0.25 * dt_dy * (E0(i, j + 1, k).z - E0(i, j, k).z + E0(i + 1, j + 1, k).z - E0(i + 1, j, k).z) + 0.5 * W.v.x * lim_x;
                corr.B.y = 0.25 * dt_dx * (E0(i + 1, j, k).z - E0(i, j, k).z + E0(i + 1, j + 1, k).z - E0(i, j + 1, k).z) + 0.5 * W.v.y * lim_y;
                uL = cadd(to_cons(zL(i, j, k)), corr); uR = cadd(to_cons(zR(i, j, k)), corr);
                uL.B.z = _B(i, j, k).z; uR.B.z = _B(i, j, k + 1).z;
                zL(i, j, k) = to_prim(uL); zR(i, j, k) = to_prim(uR);
            }
    });
}

// ------------------------------------------------------------ the steps
// Godunov::sweep + Grid1D::unsplit_step (Splitting.cpp:67-115)
static void sweep(Arr<Prim>& w, double dt_dx) {
    const int size = w.nx, g = w.g;
    auto upd = [&](int i, const Cons& d) {
        Cons U = to_cons(w(i));
        U = cadd(U, d);
        w(i) = to_prim(U);
    };
    Cons fL, fR;
    Prim LR = w(-g), RL, RR;
    muscl(w(-g), RL, w(-g + 1), RR, w(-g + 2), dt_dx, V3{});
    fL = Riemann(LR, RL).flux_X(dt_dx);
    for (int i = -g + 1; i < size + g - 1; i++) {
        LR = RR;
        if (i + 2 < size + g) muscl(w(i), RL, w(i + 1), RR, w(i + 2), dt_dx, V3{});
        else RL = w(i + 1);
        fR = Riemann(LR, RL).flux_X(dt_dx);
        upd(i, cmul(csub(fL, fR), dt_dx));
        fL = fR;
    }
    LR = RR;
    RL = w(size + g - 1);
    fR = Riemann(LR, RL).flux_X(dt_dx);
    upd(size + g - 2, cmul(csub(fL, fR), dt_dx));
}
void Grid::step1D(double dt) {
    Arr<Prim> _w = w;
    sweep(_w, dt / dx);
    for (int i = 0; i < nx; i++) {
        if (!is_physical(_w(i))) {
            char b[96]; snprintf(b, sizeof b, "Unphysical state would be produced at (%d)", i);
            throw std::runtime_error(b);
        }
    }
    w = _w;  // commits ghosts too (Splitting.cpp:114)
}

// Grid2D::unsplit_step (Unsplit.cpp:28-111)
void Grid::step2D(double dt) {
    const double dt_dx = dt / dx, dt_dy = dt / dy;
    const bool mhd = cfg.mhd;
    Arr<Prim> xL(nx, ny, g), xR(nx, ny, g), yL(nx, ny, g), yR(nx, ny, g);
    Arr<V3> Bd;
    const Arr<V3>& Bf = mhd ? B : Bd;
    half_states2D(xL, w, xR, Bf, dt_dx, dt_dy, 0);
    half_states2D(yL, w, yR, Bf, dt_dx, dt_dy, 1);
    Arr<V3> Eh;
    if (cfg.ctu) {
        if (mhd) {
            Eh = Arr<V3>(nx + 1, ny + 1, g);
            ctu_mhd2D(xL, xR, yL, yR, B, w, Eh, dt_dx, dt_dy);
        } else {
            Arr<Cons> FX(nx, ny, g), FY(nx, ny, g);
            flux2D(xL, xR, FX, 0, nx, -1, ny + 1, dt_dx, 0);
            flux2D(yL, yR, FY, -1, nx + 1, 0, ny, dt_dy, 1);
            correct_state2D(xL, xR, FY, (0.5 * dt_dy), 1);
            correct_state2D(yL, yR, FX, (0.5 * dt_dx), 0);
        }
    }
    Arr<Cons> FX(nx, ny, g), FY(nx, ny, g);
    if (mhd) {
        flux2D(xL, xR, FX, 0, nx, -1, ny + 1, dt_dx, 0);
        flux2D(yL, yR, FY, -1, nx + 1, 0, ny, dt_dy, 1);
    } else {
        flux2D(xL, xR, FX, 0, nx, 0, ny, dt_dx, 0);
        flux2D(yL, yR, FY, 0, nx, 0, ny, dt_dy, 1);
    }
    Arr<Prim> _w(nx, ny, g);
    apply_fluxes2D(w, _w, FX, FY, dt_dx, dt_dy, 0);
    Arr<V3> _B;
    if (mhd) {
        Arr<V3> E(nx + 1, ny + 1, g);
        compute_electric2D(E, FX, FY, 0);
        if (cfg.ctu) upwind_electric2D(E, FX, FY, Eh, 0);
        _B = B;
        faraday2D(E, _B, dt_dx, dt_dy, 0);
        compute_body_fields2D(_B, _w);
    }
    for (int i = 0; i < nx; i++)
        for (int j = 0; j < ny; j++)
            if (!is_physical(_w(i, j))) throw std::runtime_error(fmt2("Unphysical state would be produced at ", i, j));
    w.clone(_w, false);
    if (mhd) B.clone(_B, false);
}

// Grid3D::unsplit_step (Unsplit.cpp:115-206)
void Grid::step3D(double dt) {
    const double dt_dx = dt / dx, dt_dy = dt / dy, dt_dz = dt / dz;
    const bool mhd = cfg.mhd;
    Arr<Prim> xL(nx, ny, nz, g), xR(nx, ny, nz, g), yL(nx, ny, nz, g), yR(nx, ny, nz, g), zL(nx, ny, nz, g), zR(nx, ny, nz, g);
    Arr<V3> Bd;
    const Arr<V3>& Bf = mhd ? B : Bd;
    half_states3D(xL, w, xR, Bf, dt_dx, dt_dy, dt_dz, 0);
    half_states3D(yL, w, yR, Bf, dt_dx, dt_dy, dt_dz, 1);
    half_states3D(zL, w, zR, Bf, dt_dx, dt_dy, dt_dz, 2);
    Arr<V3> Eh;
    if (cfg.ctu) {
        if (mhd) {
            Eh = Arr<V3>(nx + 1, ny + 1, nz + 1, g);
            ctu_mhd3D(xL, xR, yL, yR, zL, zR, B, w, Eh, dt_dx, dt_dy, dt_dz);
        } else {
            ctu_hydro3D(xL, xR, yL, yR, zL, zR, dt_dx, dt_dy, dt_dz);
        }
    }
    Arr<Cons> FX(nx, ny, nz, g), FY(nx, ny, nz, g), FZ(nx, ny, nz, g);
    if (mhd) {
        flux3D(xL, xR, FX, 0, nx, -1, ny + 1, -1, nz + 1, dt_dx, 0);
        flux3D(yL, yR, FY, -1, nx + 1, 0, ny, -1, nz + 1, dt_dy, 1);
        flux3D(zL, zR, FZ, -1, nx + 1, -1, ny + 1, 0, nz, dt_dz, 2);
    } else {
        flux3D(xL, xR, FX, 0, nx, 0, ny, 0, nz, dt_dx, 0);
        flux3D(yL, yR, FY, 0, nx, 0, ny, 0, nz, dt_dy, 1);
        flux3D(zL, zR, FZ, 0, nx, 0, ny, 0, nz, dt_dz, 2);
    }
    { Arr<Prim> t0; xL = t0; xR = t0; yL = t0; yR = t0; zL = t0; zR = t0; }
    Arr<Prim> _w(nx, ny, nz, g);
    apply_fluxes3D(w, _w, FX, FY, FZ, dt_dx, dt_dy, dt_dz, 0);
    Arr<V3> _B;
    if (mhd) {
        Arr<V3> E(nx + 1, ny + 1, nz + 1, g);
        compute_electric3D(E, FX, FY, FZ, 0);
        if (cfg.ctu) upwind_electric3D(E, FX, FY, FZ, Eh, 0);
        _B = B;
        faraday3D(E, _B, dt_dx, dt_dy, dt_dz, 0);
        compute_body_fields3D(_B, _w);
    }
    for (int i = 0; i < nx; i++)
        for (int j = 0; j < ny; j++)
            for (int k = 0; k < nz; k++)
                if (!is_physical(_w(i, j, k))) throw std::runtime_error(fmt3("Unphysical state would be produced at ", i, j, k));
    w.clone(_w, false);
    if (mhd) B.clone(_B, false);
}

void Grid::unsplit_step(double dt) {
    CFG = cfg;
    if (dim == 1) step1D(dt);
    else if (dim == 2) step2D(dt);
    else step3D(dt);
}

// Grid2D/3D::advanceX/Y/Z (Splitting.cpp:219-299): the boundary pass, then
// Godunov::sweep along every line of the axis, ghost lines included; the y and
// z lines are swept in their own frame (XY / XZ swap in and out).
void Grid::advance_axis(int ax, double dt) {
    apply_boundary();
    const int n = ax == 0 ? nx : ax == 1 ? ny : nz;
    const double d = ax == 0 ? dx : ax == 1 ? dy : dz;
    const int t0n = ax == 0 ? ny : nx;                        // first transverse dim (parallel loop)
    const int t1n = dim == 3 ? (ax == 2 ? ny : nz) : 1;       // second transverse dim (3D only)
    const int t1g = dim == 3 ? g : 0;
    auto at = [&](int a, int p, int q) -> Prim& {             // a along ax, (p, q) transverse
        if (dim == 2) return ax == 0 ? w(a, p) : w(p, a);
        if (ax == 0) return w(a, p, q);
        if (ax == 1) return w(p, a, q);
        return w(p, q, a);
    };
    auto in = [&](const Prim& x) { return ax == 0 ? x : ax == 1 ? swapXY(x) : swapXZ(x); };
    pfor(-g, t0n + g, [&](int p) {
        Arr<Prim> line(n, g);
        for (int q = -t1g; q < t1n + t1g; q++) {
            for (int a = -g; a < n + g; a++) line(a) = in(at(a, p, q));
            sweep(line, dt / d);
            for (int a = -g; a < n + g; a++) at(a, p, q) = in(line(a));  // the swaps are their own inverses
        }
    });
}
// Grid2D/3D::split_step (Splitting.cpp:118-216): Strang splitting, the sweep
// order rotating with sweep_step; nothing is committed if a sweep throws.
// Grid1D::split_step is unsplit_step (Splitting.cpp:98).
void Grid::split_step(double dt) {
    CFG = cfg;
    if (dim == 1) { step1D(dt); return; }
    if (cfg.mhd) throw std::runtime_error("MHD requires Unsplit dimensions. Please enable DIMENSION_UNSPLIT in Config.h");
    const Arr<Prim> saved = w;
    const double h = dt / 2;
    try {
        if (dim == 2) {
            if (sweep_step % 2 == 0) { advance_axis(0, h); advance_axis(1, dt); advance_axis(0, h); }
            else { advance_axis(1, h); advance_axis(0, dt); advance_axis(1, h); }
        } else {
            static const int order[6][3] = {{0, 1, 2}, {2, 0, 1}, {1, 2, 0}, {2, 1, 0}, {0, 2, 1}, {1, 0, 2}};
            const int* o = order[sweep_step % 6];
            advance_axis(o[0], h); advance_axis(o[1], h); advance_axis(o[2], dt); advance_axis(o[1], h); advance_axis(o[0], h);
        }
    } catch (...) {
        w = saved;
        throw;
    }
    sweep_step++;
}

// Grid::advance_unsplit / advance_split (CFL.cpp:160-238)
long Grid::advance(double dt, bool check_cfl, long* failures, long max_steps, bool split) {
    long steps = 0, fails = 0;
    while (dt > cfg.timestep_tol) {
        if (max_steps >= 0 && steps >= max_steps) break;
        apply_boundary();
        double t1 = check_cfl ? std::min(dt, cfl_time()) : dt;
        while (true) {
            try {
                if (split) split_step(t1);
                else unsplit_step(t1);
                break;
            } catch (const std::exception&) {
                fails++;
            }
            t1 *= 0.5;
            if (t1 < cfg.timestep_tol) {
                if (failures) *failures = fails;
                throw std::runtime_error("Timestep has fallen below minimum. Exiting");
            }
        }
        dt -= t1;
        steps++;
    }
    if (failures) *failures = fails;
    return steps;
}

}  // namespace oracle

// ---------------------------------------------------- test entry points
// CT pieces on caller arrays in ExtendedArray layout (8 doubles per cell /
// flux record, 3 per face/edge), for the function-level checks.
namespace oracle {
template <class T>
static void load_arr(Arr<T>& A, const double* p) { std::memcpy(A.a.data(), p, A.a.size() * sizeof(T)); }
template <class T>
static void save_arr(const Arr<T>& A, double* p) { std::memcpy(p, A.a.data(), A.a.size() * sizeof(T)); }
static_assert(sizeof(Prim) == 64 && sizeof(Cons) == 64 && sizeof(V3) == 24, "record layout");

void test_electric(int dim, int nx, int ny, int nz, int g, const double* FX, const double* FY, const double* FZ,
                   const double* ref, int gg, double* E) {
    if (dim == 2) {
        Arr<Cons> fx(nx, ny, g), fy(nx, ny, g);
        load_arr(fx, FX); load_arr(fy, FY);
        Arr<V3> e(nx + 1, ny + 1, g);
        load_arr(e, E);
        compute_electric2D(e, fx, fy, gg);
        if (ref) { Arr<V3> r(nx + 1, ny + 1, g); load_arr(r, ref); upwind_electric2D(e, fx, fy, r, gg); }
        save_arr(e, E);
    } else {
        Arr<Cons> fx(nx, ny, nz, g), fy(nx, ny, nz, g), fz(nx, ny, nz, g);
        load_arr(fx, FX); load_arr(fy, FY); load_arr(fz, FZ);
        Arr<V3> e(nx + 1, ny + 1, nz + 1, g);
        load_arr(e, E);
        compute_electric3D(e, fx, fy, fz, gg);
        if (ref) { Arr<V3> r(nx + 1, ny + 1, nz + 1, g); load_arr(r, ref); upwind_electric3D(e, fx, fy, fz, r, gg); }
        save_arr(e, E);
    }
}
void test_body_electric(int dim, int nx, int ny, int nz, int g, const double* W, int gg, double* E) {
    if (dim == 2) {
        Arr<Prim> w(nx, ny, g); load_arr(w, W);
        Arr<V3> e(nx + 1, ny + 1, g); load_arr(e, E);
        body_electric2D(w, e, gg); save_arr(e, E);
    } else {
        Arr<Prim> w(nx, ny, nz, g); load_arr(w, W);
        Arr<V3> e(nx + 1, ny + 1, nz + 1, g); load_arr(e, E);
        body_electric3D(w, e, gg); save_arr(e, E);
    }
}
void test_faraday(int dim, int nx, int ny, int nz, int g, const double* E, double dx, double dy, double dz, int gg, double* B) {
    if (dim == 2) {
        Arr<V3> e(nx + 1, ny + 1, g), b(nx + 1, ny + 1, g);
        load_arr(e, E); load_arr(b, B);
        faraday2D(e, b, dx, dy, gg); save_arr(b, B);
    } else {
        Arr<V3> e(nx + 1, ny + 1, nz + 1, g), b(nx + 1, ny + 1, nz + 1, g);
        load_arr(e, E); load_arr(b, B);
        faraday3D(e, b, dx, dy, dz, gg); save_arr(b, B);
    }
}
void test_body_fields(int dim, int nx, int ny, int nz, int g, const double* B, double* W) {
    if (dim == 2) {
        Arr<V3> b(nx + 1, ny + 1, g); load_arr(b, B);
        Arr<Prim> w(nx, ny, g); load_arr(w, W);
        compute_body_fields2D(b, w); save_arr(w, W);
    } else {
        Arr<V3> b(nx + 1, ny + 1, nz + 1, g); load_arr(b, B);
        Arr<Prim> w(nx, ny, nz, g); load_arr(w, W);
        compute_body_fields3D(b, w); save_arr(w, W);
    }
}
// correctState (CTU.cpp:264-301) on loaded arrays; L, R updated in place.
void test_correct_state(int dim, int nx, int ny, int nz, int g, double* L, double* R, const double* F, double dt_dL,
                        int sdim, int fdim) {
    if (dim == 2) {
        Arr<Prim> l(nx, ny, g), r(nx, ny, g);
        Arr<Cons> f(nx, ny, g);
        load_arr(l, L); load_arr(r, R); load_arr(f, F);
        correct_state2D(l, r, f, dt_dL, fdim);
        save_arr(l, L); save_arr(r, R);
    } else {
        Arr<Prim> l(nx, ny, nz, g), r(nx, ny, nz, g);
        Arr<Cons> f(nx, ny, nz, g);
        load_arr(l, L); load_arr(r, R); load_arr(f, F);
        correct_state3D(l, r, f, dt_dL, sdim, fdim);
        save_arr(l, L); save_arr(r, R);
    }
}
// computeCTUFlux_X/Y/Z (CTU.cpp:306-370) on loaded arrays; F updated in place.
void test_ctu_flux(int axis, int nx, int ny, int nz, int g, const double* L, const double* R, const double* T, double* F,
                   double dt_d, double dt_t, int dimt) {
    Arr<Prim> l(nx, ny, nz, g), r(nx, ny, nz, g);
    Arr<Cons> t(nx, ny, nz, g), f(nx, ny, nz, g);
    load_arr(l, L); load_arr(r, R); load_arr(t, T); load_arr(f, F);
    ctu_flux3D(l, r, t, f, dt_d, dt_t, axis, dimt);
    save_arr(f, F);
}
}  // namespace oracle
namespace oracle {
double test_cfl_speed(const Prim& W, double dx, double dy, double dz) { return cfl_speed(W, dx, dy, dz); }
}  // namespace oracle
