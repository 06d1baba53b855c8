// oracle.hpp -- CPU restatement of DRAGON's unsplit Godunov step.
//
// TEST INFRASTRUCTURE ONLY.  This is the parity oracle for the HIP path in
// dragon_amd/.  Only tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg may load it, and only as the checker or as the timed CPU
// baseline -- never as the product path.
//
// It restates, in our own code, the reference algorithm and its operation
// order (files cited per function in oracle.cpp).  It keeps the reference's
// AoS, multi-pass structure (one loop nest per reference loop nest) so that
// it doubles as the reference-equivalent CPU baseline.  Every #ifdef of the
// reference's Config.h becomes a runtime field of OConfig.
//
// Build: oracle/Makefile (g++ -O3 -ffp-contract=off, no fast-math).
#pragma once
#include <cmath>
#include <cstdint>
#include <string>
#include <vector>
#include <stdexcept>

namespace oracle {

// ---------------------------------------------------------------- config
// Config.h:27-110 switches, as runtime values.
enum Solver { RS_EXACT = 0, RS_HLL = 1, RS_HLLC = 2, RS_HLLD = 3, RS_HLLE = 4, RS_HLLX = 5 };
enum Limiter { LIM_MINMOD = 0, LIM_MC = 1, LIM_VANLEER = 2, LIM_SUPERBEE = 3, LIM_VANALBADA = 4 };
enum CTEnergy { CT_TOTAL_E = 0, CT_THERMAL = 1, CT_BETA_GATED = 2 };
enum CFLMode { CFL_MAX = 0, CFL_ADD = 1 };  // p>1 => power mean

struct OConfig {
    int mhd = 1;
    double gamma = 5.0 / 3.0;
    int solver = RS_HLLX;
    int limiter = LIM_MC;
    int muscl = 1;              // MUSCL_Hancock
    int ctu = 1;                // CTU
    int ct_energy = CT_BETA_GATED;
    double ct_beta = 0.1;       // CT_Energy_Beta
    int cfl_mode = CFL_ADD;
    double cfl_coeff = 0.3;
    int verify_fallback = -1;   // -1: reference default (on iff hydro)
    int fallback_try_hlle = 0;  // RIEMANN_FALLBACK_TRY_HLLE
    double fallback_param = 1.0;
    int hlld_safety = 1;        // HLLD_PHYSICAL_SAFETY
    int exact_rarefaction_check = 1;
    double exact_tol = 1e-12;
    int exact_max_iters = 6;
    double timestep_tol = 1e-14;
    int threads = 1;
    bool verify() const { return verify_fallback < 0 ? !mhd : verify_fallback != 0; }
};

extern OConfig CFG;  // active configuration (set by the API before each call)

// ----------------------------------------------------------------- state
struct V3 {
    double x = 0, y = 0, z = 0;
};
// PrimitiveState (FluidElement.hpp:34-43).  B is carried but ignored in hydro.
struct Prim {
    double rho = 0, p = 0;
    V3 v, B;
};
// ConservativeState (FluidElement.hpp:70-80).
struct Cons {
    double rho = 0, E = 0;
    V3 mom, B;
};

Cons to_cons(const Prim& w);
Prim to_prim(const Cons& u);
double energy(const Prim& w);
double pressure(const Cons& u);
double enthalpy(const Prim& w);
double cs(const Prim& w);
double c_alfven(const Prim& w);
double c_fast(const Prim& w, double Bk);
double c_fast_x(const Prim& w);
double c_fast_max(const Prim& w);
bool is_physical(const Prim& w);
bool is_finite(const Cons& u);
bool is_physical(const Cons& u);
Cons flux_v(const Cons& u, const V3& v);
Cons flux_cons(const Cons& u);
Cons flux_prim(const Prim& w);
Prim swapXY(Prim w);
Prim swapXZ(Prim w);
Cons swapXY(Cons u);
Cons swapXZ(Cons u);

// limiters (Limiters.cpp)
double minmod(double a, double b);
double lim_mc(double a, double b);
double lim_vanleer(double a, double b);
double lim_superbee(double a, double b);
double lim_vanalbada(double a, double b);
double limit1(double a, double b);
Prim limit(const Prim& a, const Prim& b);
// MUSCL-Hancock (MUSCL.cpp:16-61)
void muscl(const Prim& wL, Prim& L, const Prim& wC, Prim& R, const Prim& wR, double dt_dL, const V3& dB);

// Riemann solvers.  Each throws std::runtime_error with the reference's text.
struct Riemann {
    Prim L, R;
    Riemann(const Prim& l, const Prim& r) : L(l), R(r) {}
    Cons flux(double dt_dx = 0);
    Cons flux_X(double dt_dx = 0) { return flux(dt_dx); }
    Cons flux_Y(double dt_dy = 0);
    Cons flux_Z(double dt_dz = 0);
    Cons HLL();
    Cons HLL(double sl, double sr);
    Cons HLLE();
    Cons HLLC();
    Cons HLLC(double sl, double sr);
    Cons HLLD();
    Cons HLLD_zero_B(double SL, double SR);
    Cons exact_flux();
    // exact solver internals (Exact.cpp), exposed for known-answer tests
    double exact_StarP(double pGuess);
    double exact_StarV(double pStar);
    static double f(double p, const Prim& w);
    void exact_solution(Prim& sL, Prim& sR, bool& trrs);
  private:
    void verify_and_fallback(Cons& flux, double dt_dx);
};
Prim exact_sample(const Prim& wL, const Prim& wR, const Prim& sL, const Prim& sR, double x_t);
Prim roe_avg(const Prim& L, const Prim& R);

// ---------------------------------------------------------------- arrays
// ExtendedArray{1,2,3}D (ExtendedArray.hpp): ghost-offset storage, last
// index fastest.
template <class T>
struct Arr {
    int dim = 1, nx = 0, ny = 1, nz = 1, g = 0;
    long sx = 0, sy = 0;  // strides of i and j
    std::vector<T> a;
    Arr() = default;
    Arr(int nx_, int g_) { init(1, nx_, 1, 1, g_); }
    Arr(int nx_, int ny_, int g_) { init(2, nx_, ny_, 1, g_); }
    Arr(int nx_, int ny_, int nz_, int g_) { init(3, nx_, ny_, nz_, g_); }
    void init(int d, int x, int y, int z, int gg) {
        dim = d; nx = x; ny = y; nz = z; g = gg;
        long ey = d >= 2 ? ny + 2 * g : 1, ez = d >= 3 ? nz + 2 * g : 1;
        sy = ez; sx = ey * ez;
        a.assign((size_t)(nx + 2 * g) * ey * ez, T());
    }
    long off(int i) const { return i + g; }
    long off(int i, int j) const { return (long)(i + g) * sx + (j + g); }
    long off(int i, int j, int k) const { return (long)(i + g) * sx + (long)(j + g) * sy + (k + g); }
    T& operator()(int i) { return a[off(i)]; }
    T& operator()(int i, int j) { return a[off(i, j)]; }
    T& operator()(int i, int j, int k) { return a[off(i, j, k)]; }
    const T& operator()(int i) const { return a[off(i)]; }
    const T& operator()(int i, int j) const { return a[off(i, j)]; }
    const T& operator()(int i, int j, int k) const { return a[off(i, j, k)]; }
    // clone(ref, copyghosts) (ExtendedArray.hpp:45-50, 91-98, 143-152)
    void clone(const Arr& r, bool ghosts = true);
};

// ------------------------------------------------------------ boundaries
// Boundary.hpp:139-149 face bits.
enum { XN = 1, XP = 2, YN = 4, YP = 8, ZN = 16, ZP = 32, FX = 3, FY = 12, FZ = 48, FALL = 63 };
enum BCKind { BC_OUTFLOW = 0, BC_PERIODIC = 1, BC_REFLECTIVE = 2, BC_FIXED = 3, BC_IGNORE = 4 };
struct BC {
    int kind = BC_OUTFLOW;
    int faces = FALL;
    bool corners = true;
    bool gated = false;
    Prim state;  // Fixed
};
int face_mask(const std::string& s);
BC make_bc(int kind, int faces, bool corners = true, bool gated = false, const Prim* st = nullptr);

// ------------------------------------------------------------------ grid
struct Grid {
    int dim = 1, nx = 1, ny = 1, nz = 1, g = 2;
    double dx = 1, dy = 1, dz = 1;
    OConfig cfg;
    Arr<Prim> w;
    Arr<V3> B;  // face fields, MHD 2D/3D only
    std::vector<BC> boundary;  // BoundaryList elements in order (implicit outflow added)

    Grid(int dim, int nx, int ny, int nz, double dx, double dy, double dz, int ghosts, const OConfig& c);
    int valid_ghosts(int gg) const;
    void apply_boundary();
    double cfl_time() const;
    void unsplit_step(double dt);   // throws std::runtime_error on failure (no commit)
    // Strang split step (Splitting.cpp:118-216, hydro only); sweep_step picks the
    // order of the directional sweeps and advances only on success.
    void split_step(double dt);
    int sweep_step = 0;
    // advance_unsplit / advance_split (CFL.cpp:160-238).  Returns the number of
    // accepted steps; throws if the timestep falls below tolerance (the reference
    // exit(1)s).
    long advance(double dt, bool check_cfl, long* failures = nullptr, long max_steps = -1, bool split = false);
    void initialize_B_fields();     // B_fields.cpp:117-149
  private:
    void step1D(double dt);
    void advance_axis(int ax, double dt);  // advanceX/Y/Z (Splitting.cpp:219-299)
    void step2D(double dt);
    void step3D(double dt);
};

// CT pieces exposed for tests (Electric.cpp, B_fields.cpp)
void faraday2D(const Arr<V3>& E, Arr<V3>& B, double dt_dx, double dt_dy, int gg);
void faraday3D(const Arr<V3>& E, Arr<V3>& B, double dt_dx, double dt_dy, double dt_dz, int gg);
void compute_body_fields2D(const Arr<V3>& B, Arr<Prim>& w);
void compute_body_fields3D(const Arr<V3>& B, Arr<Prim>& w);
bool should_protect_thermal(const Prim& w);
double upwind_corr(double m, double fm, double fp, double rm, double rp);

}  // namespace oracle

namespace oracle {
// Function-level entry points for the checks against the reference's own code.
void test_electric(int dim, int nx, int ny, int nz, int g, const double* FX, const double* FY, const double* FZ,
                   const double* ref, int gg, double* E);
void test_body_electric(int dim, int nx, int ny, int nz, int g, const double* W, int gg, double* E);
void test_faraday(int dim, int nx, int ny, int nz, int g, const double* E, double dx, double dy, double dz, int gg, double* B);
void test_body_fields(int dim, int nx, int ny, int nz, int g, const double* B, double* W);
void test_correct_state(int dim, int nx, int ny, int nz, int g, double* L, double* R, const double* F, double dt_dL,
                        int sdim, int fdim);
void test_ctu_flux(int axis, int nx, int ny, int nz, int g, const double* L, const double* R, const double* T, double* F,
                   double dt_d, double dt_t, int dimt);
}  // namespace oracle
namespace oracle {
double test_cfl_speed(const Prim& W, double dx, double dy, double dz);
}  // namespace oracle
