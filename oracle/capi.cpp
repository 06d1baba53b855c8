// capi.cpp -- C entry points of the oracle for the Python test harness.
// TEST INFRASTRUCTURE ONLY (see oracle.hpp).
//
// State arrays cross this boundary in the reference's ExtendedArray layout
// (last index fastest, ghosts included), 8 doubles per cell
// (rho, p, vx, vy, vz, Bx, By, Bz) and 3 doubles per face (Bx, By, Bz).
#include <cstring>

#include "oracle.hpp"

using namespace oracle;

extern "C" {

// Plain mirror of OConfig for ctypes.
struct oc_cfg {
    int mhd;
    double gamma;
    int solver, limiter, muscl, ctu, ct_energy;
    double ct_beta;
    int cfl_mode;
    double cfl_coeff;
    int verify_fallback, fallback_try_hlle;
    double fallback_param;
    int hlld_safety;
    int threads;
};

static OConfig to_cfg(const oc_cfg* c) {
    OConfig o;
    if (!c) return o;
    o.mhd = c->mhd; o.gamma = c->gamma; o.solver = c->solver; o.limiter = c->limiter;
    o.muscl = c->muscl; o.ctu = c->ctu; o.ct_energy = c->ct_energy; o.ct_beta = c->ct_beta;
    o.cfl_mode = c->cfl_mode; o.cfl_coeff = c->cfl_coeff; o.verify_fallback = c->verify_fallback;
    o.fallback_try_hlle = c->fallback_try_hlle; o.fallback_param = c->fallback_param;
    o.hlld_safety = c->hlld_safety; o.threads = c->threads < 1 ? 1 : c->threads;
    return o;
}
static void put_err(char* err, int n, const char* msg) {
    if (err && n > 0) { strncpy(err, msg, n - 1); err[n - 1] = 0; }
}
static Prim P8(const double* a) {
    Prim w; w.rho = a[0]; w.p = a[1]; w.v = {a[2], a[3], a[4]}; w.B = {a[5], a[6], a[7]}; return w;
}
static void W8(const Prim& w, double* a) {
    a[0] = w.rho; a[1] = w.p; a[2] = w.v.x; a[3] = w.v.y; a[4] = w.v.z; a[5] = w.B.x; a[6] = w.B.y; a[7] = w.B.z;
}
static Cons C8(const double* a) {
    Cons u; u.rho = a[0]; u.E = a[1]; u.mom = {a[2], a[3], a[4]}; u.B = {a[5], a[6], a[7]}; return u;
}
static void U8(const Cons& u, double* a) {
    a[0] = u.rho; a[1] = u.E; a[2] = u.mom.x; a[3] = u.mom.y; a[4] = u.mom.z; a[5] = u.B.x; a[6] = u.B.y; a[7] = u.B.z;
}

// ------------------------------------------------------------------ grid
void* oc_grid_create(int dim, int nx, int ny, int nz, double dx, double dy, double dz, int ghosts, const oc_cfg* c) {
    return new Grid(dim, nx, ny, nz, dx, dy, dz, ghosts, to_cfg(c));
}
void oc_grid_destroy(void* h) { delete (Grid*)h; }
int oc_grid_ghosts(void* h) { return ((Grid*)h)->g; }
long oc_grid_ncells(void* h) { return (long)((Grid*)h)->w.a.size(); }
long oc_grid_nfaces(void* h) { return (long)((Grid*)h)->B.a.size(); }
void oc_grid_set_threads(void* h, int t) { ((Grid*)h)->cfg.threads = t < 1 ? 1 : t; }
void oc_grid_get_w(void* h, double* out) {
    auto& w = ((Grid*)h)->w.a;
    for (size_t i = 0; i < w.size(); i++) W8(w[i], out + 8 * i);
}
void oc_grid_set_w(void* h, const double* in) {
    auto& w = ((Grid*)h)->w.a;
    for (size_t i = 0; i < w.size(); i++) w[i] = P8(in + 8 * i);
}
void oc_grid_get_B(void* h, double* out) {
    auto& B = ((Grid*)h)->B.a;
    for (size_t i = 0; i < B.size(); i++) { out[3 * i] = B[i].x; out[3 * i + 1] = B[i].y; out[3 * i + 2] = B[i].z; }
}
void oc_grid_set_B(void* h, const double* in) {
    auto& B = ((Grid*)h)->B.a;
    for (size_t i = 0; i < B.size(); i++) B[i] = {in[3 * i], in[3 * i + 1], in[3 * i + 2]};
}
void oc_grid_clear_bc(void* h) { ((Grid*)h)->boundary.clear(); }
void oc_grid_add_bc(void* h, int kind, int faces, int corners, int gated, const double* st8) {
    Prim s;
    if (st8) s = P8(st8);
    ((Grid*)h)->boundary.push_back(make_bc(kind, faces, corners != 0, gated != 0, st8 ? &s : nullptr));
}
void oc_grid_apply_bc(void* h) { ((Grid*)h)->apply_boundary(); }
void oc_grid_init_B(void* h) { CFG = ((Grid*)h)->cfg; ((Grid*)h)->initialize_B_fields(); }
int oc_grid_cfl(void* h, double* dt, char* err, int n) {
    try { *dt = ((Grid*)h)->cfl_time(); return 0; } catch (const std::exception& e) { put_err(err, n, e.what()); return 1; }
}
int oc_grid_step(void* h, double dt, char* err, int n) {
    try { ((Grid*)h)->unsplit_step(dt); return 0; } catch (const std::exception& e) { put_err(err, n, e.what()); return 1; }
}
// Returns accepted steps, or -1 on an unrecoverable failure (message in err).
long oc_grid_advance(void* h, double dt, int check_cfl, long max_steps, long* failures, char* err, int n) {
    try {
        return ((Grid*)h)->advance(dt, check_cfl != 0, failures, max_steps);
    } catch (const std::exception& e) { put_err(err, n, e.what()); return -1; }
}
int oc_grid_split_step(void* h, double dt, char* err, int n) {
    try {
        ((Grid*)h)->split_step(dt);
        return 0;
    } catch (const std::exception& e) { put_err(err, n, e.what()); return 1; }
}
long oc_grid_advance_split(void* h, double dt, int check_cfl, long max_steps, long* failures, char* err, int n) {
    try {
        return ((Grid*)h)->advance(dt, check_cfl != 0, failures, max_steps, true);
    } catch (const std::exception& e) { put_err(err, n, e.what()); return -1; }
}
// Faraday B = curl(A) initialisation (main.cpp:35-43, 57-67); A has no ghosts,
// size (nx+1)(ny+1)[(nz+1)], 3 doubles per corner.
void oc_grid_faraday_init(void* h, const double* A3) {
    Grid& G = *(Grid*)h;
    CFG = G.cfg;
    if (G.dim == 2) {
        Arr<V3> A(G.nx + 1, G.ny + 1, 0);
        for (size_t i = 0; i < A.a.size(); i++) A.a[i] = {A3[3 * i], A3[3 * i + 1], A3[3 * i + 2]};
        for (int i = 0; i <= G.nx; i++) for (int j = 0; j <= G.ny; j++) G.B(i, j) = V3{};
        faraday2D(A, G.B, -1 / G.dx, -1 / G.dy, 0);
    } else if (G.dim == 3) {
        Arr<V3> A(G.nx + 1, G.ny + 1, G.nz + 1, 0);
        for (size_t i = 0; i < A.a.size(); i++) A.a[i] = {A3[3 * i], A3[3 * i + 1], A3[3 * i + 2]};
        for (int i = 0; i <= G.nx; i++) for (int j = 0; j <= G.ny; j++) for (int k = 0; k <= G.nz; k++) G.B(i, j, k) = V3{};
        faraday3D(A, G.B, -1 / G.dx, -1 / G.dy, -1 / G.dz, 0);
    }
}

// --------------------------------------------------------- per-function
// dir: 0 x, 1 y, 2 z.  Returns 0, or 1 with the reference's message.
int oc_riemann_flux(const oc_cfg* c, const double* L8, const double* R8, double dt_dx, int dir, double* F8, char* err, int n) {
    CFG = to_cfg(c);
    try {
        Riemann r(P8(L8), P8(R8));
        Cons F = dir == 0 ? r.flux_X(dt_dx) : dir == 1 ? r.flux_Y(dt_dx) : r.flux_Z(dt_dx);
        U8(F, F8);
        return 0;
    } catch (const std::exception& e) { put_err(err, n, e.what()); return 1; }
}
// which: 0 HLL, 1 HLLE, 2 HLLC, 3 HLLD, 4 exact, 5 HLL(sl,sr), 6 HLLC(sl,sr)
void oc_riemann_solver(const oc_cfg* c, const double* L8, const double* R8, int which, double sl, double sr, double* F8) {
    CFG = to_cfg(c);
    Riemann r(P8(L8), P8(R8));
    Cons F;
    switch (which) {
        case 0: F = r.HLL(); break;
        case 1: F = r.HLLE(); break;
        case 2: F = r.HLLC(); break;
        case 3: F = r.HLLD(); break;
        case 4: F = r.exact_flux(); break;
        case 5: F = r.HLL(sl, sr); break;
        default: F = r.HLLC(sl, sr); break;
    }
    U8(F, F8);
}
// Exact star state: out = {p*, u*, rho*L, rho*R, trrs}
void oc_exact_star(const oc_cfg* c, const double* L8, const double* R8, double* out) {
    CFG = to_cfg(c);
    Riemann r(P8(L8), P8(R8));
    Prim sL, sR;
    bool t;
    r.exact_solution(sL, sR, t);
    out[0] = sL.p; out[1] = sL.v.x; out[2] = sL.rho; out[3] = sR.rho; out[4] = t ? 1 : 0;
}
void oc_exact_sample(const oc_cfg* c, const double* L8, const double* R8, double x_t, double* W8o) {
    CFG = to_cfg(c);
    Riemann r(P8(L8), P8(R8));
    Prim sL, sR;
    bool t;
    r.exact_solution(sL, sR, t);
    W8(exact_sample(r.L, r.R, sL, sR, x_t), W8o);
}
double oc_exact_f(const oc_cfg* c, double p, const double* w8) { CFG = to_cfg(c); return Riemann::f(p, P8(w8)); }
void oc_muscl(const oc_cfg* c, const double* wL, const double* wC, const double* wR, double dt_dL, const double* dB3, double* L8, double* R8) {
    CFG = to_cfg(c);
    Prim l, r;
    V3 dB;
    if (dB3) dB = {dB3[0], dB3[1], dB3[2]};
    muscl(P8(wL), l, P8(wC), r, P8(wR), dt_dL, dB);
    W8(l, L8); W8(r, R8);
}
double oc_limit(const oc_cfg* c, double a, double b) { CFG = to_cfg(c); return limit1(a, b); }
void oc_to_cons(const oc_cfg* c, const double* w8, double* u8) { CFG = to_cfg(c); U8(to_cons(P8(w8)), u8); }
void oc_to_prim(const oc_cfg* c, const double* u8, double* w8) { CFG = to_cfg(c); W8(to_prim(C8(u8)), w8); }
void oc_flux_prim(const oc_cfg* c, const double* w8, double* F8) { CFG = to_cfg(c); U8(flux_prim(P8(w8)), F8); }
// out = {cs, c_alfven, c_fast_x, c_fast_max, enthalpy, energy}
void oc_speeds(const oc_cfg* c, const double* w8, double* out) {
    CFG = to_cfg(c);
    Prim w = P8(w8);
    out[0] = cs(w);
    out[1] = CFG.mhd ? c_alfven(w) : 0;
    out[2] = CFG.mhd ? c_fast_x(w) : 0;
    out[3] = CFG.mhd ? c_fast_max(w) : 0;
    out[4] = enthalpy(w);
    out[5] = energy(w);
}
int oc_is_physical(const oc_cfg* c, const double* w8) { CFG = to_cfg(c); return is_physical(P8(w8)) ? 1 : 0; }
double oc_upwind_corr(double m, double fm, double fp, double rm, double rp) { return upwind_corr(m, fm, fp, rm, rp); }
int oc_protect_thermal(const oc_cfg* c, const double* w8) { CFG = to_cfg(c); return should_protect_thermal(P8(w8)) ? 1 : 0; }
int oc_face_mask(const char* s) { return face_mask(s); }
void oc_ct_electric(const oc_cfg* c, int dim, int nx, int ny, int nz, int g, const double* FX, const double* FY, const double* FZ,
                    const double* ref, int gg, double* E) {
    CFG = to_cfg(c);
    test_electric(dim, nx, ny, nz, g, FX, FY, FZ, ref, gg, E);
}
void oc_ct_body_electric(const oc_cfg* c, int dim, int nx, int ny, int nz, int g, const double* W, int gg, double* E) {
    CFG = to_cfg(c);
    test_body_electric(dim, nx, ny, nz, g, W, gg, E);
}
void oc_ct_faraday(const oc_cfg* c, int dim, int nx, int ny, int nz, int g, const double* E, double dx, double dy, double dz, int gg,
                   double* B) {
    CFG = to_cfg(c);
    test_faraday(dim, nx, ny, nz, g, E, dx, dy, dz, gg, B);
}
void oc_ct_body_fields(const oc_cfg* c, int dim, int nx, int ny, int nz, int g, const double* B, double* W) {
    CFG = to_cfg(c);
    test_body_fields(dim, nx, ny, nz, g, B, W);
}
void oc_correct_state(const oc_cfg* c, int dim, int nx, int ny, int nz, int g, double* L, double* R, const double* F,
                      double dt_dL, int sdim, int fdim) {
    CFG = to_cfg(c);
    test_correct_state(dim, nx, ny, nz, g, L, R, F, dt_dL, sdim, fdim);
}
int oc_ctu_flux(const oc_cfg* c, int axis, int nx, int ny, int nz, int g, const double* L, const double* R, const double* T,
                double* F, double dt_d, double dt_t, int dimt, char* err, int n) {
    CFG = to_cfg(c);
    try {
        test_ctu_flux(axis, nx, ny, nz, g, L, R, T, F, dt_d, dt_t, dimt);
        return 0;
    } catch (const std::exception& e) {
        if (err && n > 0) { std::strncpy(err, e.what(), n - 1); err[n - 1] = 0; }
        return 1;
    }
}
double oc_cfl_speed(const oc_cfg* c, const double* w8, double dx, double dy, double dz) {
    CFG = to_cfg(c);
    return test_cfl_speed(P8(w8), dx, dy, dz);
}

}  // extern "C"
