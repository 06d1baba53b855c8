// ref_driver.cpp -- C entry points over the REFERENCE's own per-function code.
// TEST INFRASTRUCTURE ONLY.
//
// Compiled together with the reference's numerics sources, straight from
// /root/reference (oracle/build_ref.sh), into oracle/_ref/libref_*.so.  It is
// used only to check the oracle restatement function by function (Riemann
// solvers, MUSCL-Hancock, limiters, state algebra, CT pieces, CFL speeds).
// Nothing here is shipped or run on the product path.
//
// The reference's step orchestration (Hydro/Godunov/Unsplit.cpp,
// Splitting.cpp) includes <format>, which this image's C++ library lacks, so
// whole-step trajectories cannot be produced from the reference here.  The
// hydro build also compiles Hydro/Godunov/CTU.cpp, whose correctState and
// computeCTUFlux_X/Y/Z (the 3D 12-solve corner fluxes) need nothing from
// those two files; ctu_sweep_hydro / ctu_sweep_MHD call Unsplit.cpp's
// computeFlux_* and applyFluxes and are not exported (the linker drops them).
#include <cstring>
#include <stdexcept>

#include "ArrayTypes.hpp"
#include "CFL.hpp"
#include "CT.hpp"
#include "Config.h"
#include "FluidElement.hpp"
#include "Riemann.hpp"
#include "TVD.hpp"
#include "Constants.h"
#ifndef MHD
#include "Unsplit.hpp"
#endif

#define API extern "C" __attribute__((visibility("default")))

static PrimitiveState P8(const double* a) {
    PrimitiveState w;
    w.rho = a[0]; w.p = a[1]; w.v = vec3(a[2], a[3], a[4]);
#ifdef MHD
    w.B = vec3(a[5], a[6], a[7]);
#endif
    return w;
}
static void W8(const PrimitiveState& w, double* a) {
    a[0] = w.rho; a[1] = w.p; a[2] = w.v.x; a[3] = w.v.y; a[4] = w.v.z;
#ifdef MHD
    a[5] = w.B.x; a[6] = w.B.y; a[7] = w.B.z;
#else
    a[5] = a[6] = a[7] = 0;
#endif
}
static ConservativeState C8(const double* a) {
    ConservativeState u;
    u.rho = a[0]; u.E = a[1]; u.mom = vec3(a[2], a[3], a[4]);
#ifdef MHD
    u.B = vec3(a[5], a[6], a[7]);
#endif
    return u;
}
static void U8(const ConservativeState& u, double* a) {
    a[0] = u.rho; a[1] = u.E; a[2] = u.mom.x; a[3] = u.mom.y; a[4] = u.mom.z;
#ifdef MHD
    a[5] = u.B.x; a[6] = u.B.y; a[7] = u.B.z;
#else
    a[5] = a[6] = a[7] = 0;
#endif
}
static void put_err(char* err, int n, const char* msg) {
    if (err && n > 0) { strncpy(err, msg, n - 1); err[n - 1] = 0; }
}

API int rf_is_mhd() {
#ifdef MHD
    return 1;
#else
    return 0;
#endif
}
API int rf_is_testmode() {
#ifdef TESTMODE
    return 1;
#else
    return 0;
#endif
}
API double rf_gamma() { return _gamma; }
// Runtime switches exist only in TESTMODE builds (Config.h:97-110).
API void rf_set(int riemann, int limiter, int cfl, int ct) {
#ifdef TESTMODE
    CONFIG::riemann_choice = riemann;
    CONFIG::limiter_choice = limiter;
    CONFIG::cfl_choice = cfl;
    CONFIG::CT_energy_choice = ct;
#else
    (void)riemann; (void)limiter; (void)cfl; (void)ct;
#endif
}

API int rf_riemann_flux(const double* L8, const double* R8, double dt_dx, int dir, double* F8, char* err, int n) {
    try {
        Riemann r(P8(L8), P8(R8));
        ConservativeState F = dir == 0 ? r.flux_X(dt_dx) : dir == 1 ? r.flux_Y(dt_dx) : r.flux_Z(dt_dx);
        U8(F, F8);
        return 0;
    } catch (const std::exception& e) { put_err(err, n, e.what()); return 1; }
}
// which: 0 HLL, 1 HLLE, 2 HLLC, 3 HLLD, 4 exact, 5 HLL(sl,sr), 6 HLLC(sl,sr)
API void rf_riemann_solver(const double* L8, const double* R8, int which, double sl, double sr, double* F8) {
    Riemann r(P8(L8), P8(R8));
    ConservativeState F;
    switch (which) {
        case 0: F = r.HLL(); break;
        case 1: F = r.HLLE(); break;
#ifdef HYDRO_AVAILABLE
        case 2: F = r.HLLC(); break;
        case 6: F = r.HLLC(sl, sr); break;
#endif
#ifdef MHD
        case 3: F = r.HLLD(); break;
#endif
        case 4: F = r.exact().flux(); break;
        case 5: F = r.HLL(sl, sr); break;
        default: break;
    }
    U8(F, F8);
}
API void rf_exact_star(const double* L8, const double* R8, double* out) {
    Riemann r(P8(L8), P8(R8));
    RiemannSolution s = r.exact();
    out[0] = s.sL.p; out[1] = s.sL.v.x; out[2] = s.sL.rho; out[3] = s.sR.rho;
}
API void rf_exact_sample(const double* L8, const double* R8, double x_t, double* W8o) {
    Riemann r(P8(L8), P8(R8));
    RiemannSolution s = r.exact();
    W8(s.sample(x_t), W8o);
}
API double rf_exact_f(double p, const double* w8) { return Riemann::f(p, P8(w8)); }
API void rf_muscl(const double* wL, const double* wC, const double* wR, double dt_dL, const double* dB3, double* L8, double* R8) {
    PrimitiveState l, r;
    vec3 dB(0, 0, 0);
    if (dB3) dB = vec3(dB3[0], dB3[1], dB3[2]);
    TVD::MUSCL(P8(wL), l, P8(wC), r, P8(wR), dt_dL, dB);
    W8(l, L8); W8(r, R8);
}
API double rf_limit(double a, double b) {
    PrimitiveState A, Bs;
    A.rho = a; Bs.rho = b;
    return TVD::limit(A, Bs).rho;
}
API void rf_to_cons(const double* w8, double* u8) { U8(ConservativeState(P8(w8)), u8); }
API void rf_to_prim(const double* u8, double* w8) { W8(PrimitiveState(C8(u8)), w8); }
API void rf_flux_prim(const double* w8, double* F8) { U8(P8(w8).flux(), F8); }
API void rf_speeds(const double* w8, double* out) {
    PrimitiveState w = P8(w8);
    out[0] = w.cs();
#ifdef MHD
    out[1] = w.c_alfven(); out[2] = w.c_fast(); out[3] = w.c_fast_max();
#else
    out[1] = out[2] = out[3] = 0;
#endif
    out[4] = w.enthalpy();
    out[5] = w.energy();
}
API int rf_is_physical(const double* w8) { return P8(w8).isPhysical() ? 1 : 0; }
API double rf_cfl_speed(const double* w8, double dx, double dy, double dz) { return CFL::cfl_speed(P8(w8), dx, dy, dz); }

// ---- small 2D/3D arrays in ExtendedArray layout (ghosts included, last index fastest)
static void load(FluxArray2D& A, const double* a) {
    int g = A.getGhosts(), m = 0;
    for (int i = -g; i < A.getSizeX() + g; i++)
        for (int j = -g; j < A.getSizeY() + g; j++, m += 8) A[i, j] = C8(a + m);
}
static void load(FluidArray2D& A, const double* a) {
    int g = A.getGhosts(), m = 0;
    for (int i = -g; i < A.getSizeX() + g; i++)
        for (int j = -g; j < A.getSizeY() + g; j++, m += 8) A[i, j] = P8(a + m);
}
static void save(const FluidArray2D& A, double* a) {
    int g = A.getGhosts(), m = 0;
    for (int i = -g; i < A.getSizeX() + g; i++)
        for (int j = -g; j < A.getSizeY() + g; j++, m += 8) W8(A[i, j], a + m);
}
static void load(FluxArray3D& A, const double* a) {
    int g = A.getGhosts(), m = 0;
    for (int i = -g; i < A.getSizeX() + g; i++)
        for (int j = -g; j < A.getSizeY() + g; j++)
            for (int k = -g; k < A.getSizeZ() + g; k++, m += 8) A[i, j, k] = C8(a + m);
}
static void save(const FluxArray3D& A, double* a) {
    int g = A.getGhosts(), m = 0;
    for (int i = -g; i < A.getSizeX() + g; i++)
        for (int j = -g; j < A.getSizeY() + g; j++)
            for (int k = -g; k < A.getSizeZ() + g; k++, m += 8) U8(A[i, j, k], a + m);
}
static void load(FluidArray3D& A, const double* a) {
    int g = A.getGhosts(), m = 0;
    for (int i = -g; i < A.getSizeX() + g; i++)
        for (int j = -g; j < A.getSizeY() + g; j++)
            for (int k = -g; k < A.getSizeZ() + g; k++, m += 8) A[i, j, k] = P8(a + m);
}
static void save(const FluidArray3D& A, double* a) {
    int g = A.getGhosts(), m = 0;
    for (int i = -g; i < A.getSizeX() + g; i++)
        for (int j = -g; j < A.getSizeY() + g; j++)
            for (int k = -g; k < A.getSizeZ() + g; k++, m += 8) W8(A[i, j, k], a + m);
}

#ifdef MHD
static void load(MagneticArray2D& A, const double* a) {
    int g = A.getGhosts(), m = 0;
    for (int i = -g; i < A.getSizeX() + g; i++)
        for (int j = -g; j < A.getSizeY() + g; j++, m += 3) A[i, j] = vec3(a[m], a[m + 1], a[m + 2]);
}
static void save(const MagneticArray2D& A, double* a) {
    int g = A.getGhosts(), m = 0;
    for (int i = -g; i < A.getSizeX() + g; i++)
        for (int j = -g; j < A.getSizeY() + g; j++, m += 3) { a[m] = A[i, j].x; a[m + 1] = A[i, j].y; a[m + 2] = A[i, j].z; }
}
static void load(MagneticArray3D& A, const double* a) {
    int g = A.getGhosts(), m = 0;
    for (int i = -g; i < A.getSizeX() + g; i++)
        for (int j = -g; j < A.getSizeY() + g; j++)
            for (int k = -g; k < A.getSizeZ() + g; k++, m += 3) A[i, j, k] = vec3(a[m], a[m + 1], a[m + 2]);
}
static void save(const MagneticArray3D& A, double* a) {
    int g = A.getGhosts(), m = 0;
    for (int i = -g; i < A.getSizeX() + g; i++)
        for (int j = -g; j < A.getSizeY() + g; j++)
            for (int k = -g; k < A.getSizeZ() + g; k++, m += 3) { a[m] = A[i, j, k].x; a[m + 1] = A[i, j, k].y; a[m + 2] = A[i, j, k].z; }
}
// E = computeElectric(FX, FY[, FZ], gg) then, if ref != NULL,
// upwindElectric(E, FX, FY[, FZ], ref, gg).  Cell arrays are (n, g); edge
// arrays (n+1, g).
API void rf_electric2D(int nx, int ny, int g, const double* FX, const double* FY, const double* ref, int gg, double* E) {
    FluxArray2D fx(nx, ny, g), fy(nx, ny, g);
    load(fx, FX); load(fy, FY);
    MagneticArray2D e(nx + 1, ny + 1, g);
    load(e, E);
    CT::computeElectric(e, fx, fy, gg);
    if (ref) {
        MagneticArray2D r(nx + 1, ny + 1, g);
        load(r, ref);
        CT::upwindElectric(e, fx, fy, r, gg);
    }
    save(e, E);
}
API void rf_electric3D(int nx, int ny, int nz, int g, const double* FX, const double* FY, const double* FZ, const double* ref, int gg, double* E) {
    FluxArray3D fx(nx, ny, nz, g), fy(nx, ny, nz, g), fz(nx, ny, nz, g);
    load(fx, FX); load(fy, FY); load(fz, FZ);
    MagneticArray3D e(nx + 1, ny + 1, nz + 1, g);
    load(e, E);
    CT::computeElectric(e, fx, fy, fz, gg);
    if (ref) {
        MagneticArray3D r(nx + 1, ny + 1, nz + 1, g);
        load(r, ref);
        CT::upwindElectric(e, fx, fy, fz, r, gg);
    }
    save(e, E);
}
API void rf_body_electric2D(int nx, int ny, int g, const double* W, int gg, double* E) {
    FluidArray2D w(nx, ny, g);
    load(w, W);
    MagneticArray2D e(nx + 1, ny + 1, g);
    load(e, E);
    CT::bodyElectric(w, e, gg);
    save(e, E);
}
API void rf_body_electric3D(int nx, int ny, int nz, int g, const double* W, int gg, double* E) {
    FluidArray3D w(nx, ny, nz, g);
    load(w, W);
    MagneticArray3D e(nx + 1, ny + 1, nz + 1, g);
    load(e, E);
    CT::bodyElectric(w, e, gg);
    save(e, E);
}
API void rf_faraday2D(int nx, int ny, int g, const double* E, double dt_dx, double dt_dy, int gg, double* B) {
    MagneticArray2D e(nx + 1, ny + 1, g), b(nx + 1, ny + 1, g);
    load(e, E); load(b, B);
    CT::Faraday(e, b, dt_dx, dt_dy, gg);
    save(b, B);
}
API void rf_faraday3D(int nx, int ny, int nz, int g, const double* E, double dt_dx, double dt_dy, double dt_dz, int gg, double* B) {
    MagneticArray3D e(nx + 1, ny + 1, nz + 1, g), b(nx + 1, ny + 1, nz + 1, g);
    load(e, E); load(b, B);
    CT::Faraday(e, b, dt_dx, dt_dy, dt_dz, gg);
    save(b, B);
}
API void rf_body_fields2D(int nx, int ny, int g, const double* B, double* W) {
    MagneticArray2D b(nx + 1, ny + 1, g);
    load(b, B);
    FluidArray2D w(nx, ny, g);
    load(w, W);
    CT::computeBodyFields(b, w);
    save(w, W);
}
API void rf_body_fields3D(int nx, int ny, int nz, int g, const double* B, double* W) {
    MagneticArray3D b(nx + 1, ny + 1, nz + 1, g);
    load(b, B);
    FluidArray3D w(nx, ny, nz, g);
    load(w, W);
    CT::computeBodyFields(b, w);
    save(w, W);
}
#endif

#ifndef MHD
// ---- CTU.cpp: correctState 2D/3D and the 12-solve corner fluxes
// L, R: half-state arrays (n, g); F, T: flux arrays (n, g), all in place.
API void rf_correct_state2D(int nx, int ny, int g, double* L, double* R, const double* F, double dt_dL, int dim) {
    FluidArray2D l(nx, ny, g), r(nx, ny, g);
    FluxArray2D f(nx, ny, g);
    load(l, L); load(r, R); load(f, F);
    correctState(l, r, f, dt_dL, dim);
    save(l, L); save(r, R);
}
API void rf_correct_state3D(int nx, int ny, int nz, int g, double* L, double* R, const double* F, double dt_dL, int sdim, int fdim) {
    FluidArray3D l(nx, ny, nz, g), r(nx, ny, nz, g);
    FluxArray3D f(nx, ny, nz, g);
    load(l, L); load(r, R); load(f, F);
    correctState(l, r, f, dt_dL, sdim, fdim);
    save(l, L); save(r, R);
}
// computeCTUFlux_{X,Y,Z}(L, R, T, F, dt_d, dt_t, dimt) for axis 0/1/2.
API int rf_ctu_flux3D(int axis, int nx, int ny, int nz, int g, const double* L, const double* R, const double* T,
                      double* F, double dt_d, double dt_t, int dimt, char* err, int n) {
    try {
        FluidArray3D l(nx, ny, nz, g), r(nx, ny, nz, g);
        FluxArray3D t(nx, ny, nz, g), f(nx, ny, nz, g);
        load(l, L); load(r, R); load(t, T); load(f, F);
        if (axis == 0) computeCTUFlux_X(l, r, t, f, dt_d, dt_t, dimt);
        else if (axis == 1) computeCTUFlux_Y(l, r, t, f, dt_d, dt_t, dimt);
        else computeCTUFlux_Z(l, r, t, f, dt_d, dt_t, dimt);
        save(f, F);
        return 0;
    } catch (const std::exception& e) { put_err(err, n, e.what()); return 1; }
}
#endif
