// dg_api.hip -- host runtime behind include/dragon_gpu.h.
//
// One handle = one device-resident grid + one HIP stream.  The state is
// double-buffered: a step writes the complete next state (cells and faces),
// and the commit is a pointer swap made only when the device failure word is
// clear -- the all-or-nothing commit of the reference's checkpoint barrier
// (Unsplit.cpp:101-110) without any copy.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/dragon_gpu.h"
#include "dg_kernels.hpp"

using namespace dg;
using namespace dg::p64;  // the FP64 kernels; the mixed form dispatches to dg::p32 explicitly

struct dg_grid {
    dg_desc d;
    int g = 0, nf = 5;
    Ext ce, fe;
    Phys ph;
    hipStream_t s = nullptr;
    std::vector<dg_bc> bcs;
    int (*bc_cb)(void* user) = nullptr;  // boundary callback (dg_set_boundary_callback)
    void* bc_user = nullptr;
    double *W = nullptr, *Wn = nullptr;
    double *B = nullptr, *Bn = nullptr, *Bh = nullptr;
    double *HS = nullptr, *F = nullptr, *Fp = nullptr, *T = nullptr;
    bool fused = true;
    int march_chunk = 0;
    int march_window = 0;
    double* Ws = nullptr;       // split step: third state buffer (allocated on first use)
    int sweep_step = 0;         // split step: rotation of the sweep order (Grid.hpp:93, 135)
    double *E0 = nullptr, *Eh = nullptr;
    double* cflpart = nullptr;  // per-block CFL maxima of the committed state, written by the apply pass
    long ncflpart = 0;
    bool cfl_cached = false;    // cflpart describes the current W (set by a committed step)
    unsigned long long* dev_words = nullptr;   // [0] failure key, [1] CFL max bits, [2..3] diagnostics
    unsigned long long* host_words = nullptr;  // pinned mirror
    std::string err;
    // optional live timing of the step kernels with HIP events
    bool timing = false;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    double step_ms = 0;
    long timed_steps = 0;
    // per-kernel timing: an event after each launch, tagged with its kernel
    // class (-1 opens a region), resolved after the next stream sync
    std::vector<hipEvent_t> evs;
    std::vector<int> ev_id;
    int n_ev = 0;
    double kms[K_COUNT] = {};
    long klaunch[K_COUNT] = {};
    // frame snapshots (dg_snapshot): device staging in file order, a copy
    // stream for the device-to-host transfer, events ordering the two
    double* snap = nullptr;
    long snap_len = 0;
    hipStream_t cs = nullptr;
    hipEvent_t snap_ready = nullptr, snap_done = nullptr;
    bool snap_recorded = false;  // snap_done has been recorded (set by dg_snapshot only)
};

static const unsigned long long NO_ERR = ~0ull;

#define HIPCHK(x)                                                              \
    do {                                                                       \
        hipError_t e_ = (x);                                                   \
        if (e_ != hipSuccess) {                                                \
            if (g) g->err = std::string(#x) + ": " + hipGetErrorString(e_);   \
            return DG_ERR_HIP;                                                 \
        }                                                                      \
    } while (0)

static void tm_record(dg_grid* g, int id) {
    if (!g->timing) return;
    if (g->n_ev == (int)g->evs.size()) {
        hipEvent_t e;
        if (hipEventCreate(&e) != hipSuccess) return;
        g->evs.push_back(e);
        g->ev_id.push_back(-1);
    }
    if (hipEventRecord(g->evs[g->n_ev], g->s) != hipSuccess) return;
    g->ev_id[g->n_ev++] = id;
}
static void tm_mark(void* user, int id) { tm_record((dg_grid*)user, id); }
// Call only after the stream has been synchronised.
static void tm_flush(dg_grid* g) {
    for (int i = 1; i < g->n_ev; i++) {
        const int id = g->ev_id[i];
        float ms = 0;
        if (id < 0 || hipEventElapsedTime(&ms, g->evs[i - 1], g->evs[i]) != hipSuccess) continue;
        g->kms[id] += ms;
        g->klaunch[id]++;
    }
    g->n_ev = 0;
}

static int fail(dg_grid* g, int code, const std::string& m) {
    if (g) g->err = m;
    return code;
}
static void set_status(dg_status* st, int code, int i, int j, int k, const std::string& m) {
    if (!st) return;
    st->code = code; st->i = i; st->j = j; st->k = k;
    snprintf(st->msg, sizeof st->msg, "%s", m.c_str());
}

static Phys make_phys(const dg_desc& d) {
    Phys p{};
    const double g = d.gamma;
    p.gamma = g; p.gm1 = g - 1.0;
    p.k1_4pi = 0.25 / M_PI; p.k1_8pi = 0.125 / M_PI; p.sq4pi = std::sqrt(4.0 * M_PI);
    p.Gm1_2G = (g - 1.0) / (2.0 * g); p.Gp1_2G = (g + 1.0) / (2.0 * g); p._2G_Gm1 = 2.0 * g / (g - 1.0);
    p.Gm1_2 = (g - 1.0) / 2.0; p._2_Gm1 = 2.0 / (g - 1.0); p._2_Gp1 = 2.0 / (g + 1.0);
    p.Gm1_Gp1 = (g - 1.0) / (g + 1.0); p.Ginv = 1.0 / g;
    p.limiter = d.limiter; p.muscl = d.muscl; p.solver = d.solver;
    p.verify = d.verify_fallback < 0 ? !d.mhd : d.verify_fallback != 0;
    p.try_hlle = d.fallback_try_hlle; p.hlld_safety = d.hlld_safety; p.fallback_param = d.fallback_param;
    p.exact_tol = 1e-12; p.exact_iters = 6; p.exact_rare_check = 1;
    p.ct_energy = d.ct_energy; p.ct_beta = d.ct_beta;
    p.eightpi = 8 * M_PI;
    return p;
}

static Soa soa(double* p, long stride) { return Soa{p, stride}; }

static StepCtx make_ctx(dg_grid* g, double dt) {
    StepCtx c{};
    c.dim = g->d.dim; c.mhd = g->d.mhd != 0; c.mixed = g->d.precision == DG_MIXED;
    c.ctu = g->d.ctu; c.cfl_mode = g->d.cfl_mode;
    c.fused = g->fused;
    c.march_chunk = g->march_chunk;
    c.march_window = g->march_window;
    c.ce = g->ce; c.fe = g->fe; c.ph = g->ph;
    c.d[0] = g->d.dx; c.d[1] = g->d.dim >= 2 ? g->d.dy : 1.0; c.d[2] = g->d.dim >= 3 ? g->d.dz : 1.0;
    c.dt = dt;
    const long nc = g->ce.size, nfc = g->fe.size;
    c.W = soa(g->W, nc); c.Wn = soa(g->Wn, nc);
    for (int q = 0; q < 3; q++) {
        c.B[q] = g->B ? g->B + q * nfc : nullptr;
        c.Bn[q] = g->Bn ? g->Bn + q * nfc : nullptr;
        c.Bh[q] = g->Bh ? g->Bh + q * nfc : nullptr;
        c.E0[q] = g->E0 ? g->E0 + q * nfc : nullptr;
        c.Eh[q] = g->Eh ? g->Eh + q * nfc : nullptr;
    }
    if (g->HS)
        for (int q = 0; q < 2 * g->d.dim; q++) c.HS[q] = soa(g->HS + (long)q * g->nf * nc, nc);
    for (int q = 0; q < g->d.dim; q++) c.F[q] = soa(g->F + (long)q * g->nf * nc, nc);
    if (g->Fp)
        for (int q = 0; q < g->d.dim; q++) c.Fp[q] = soa(g->Fp + (long)q * g->nf * nc, nc);
    if (g->T)
        for (int q = 0; q < 5; q++) c.T[q] = soa(g->T + (long)q * g->nf * nc, nc);
    c.err = g->dev_words;
    c.cflpart = g->cflpart;
    c.mark = g->timing ? tm_mark : nullptr;
    c.mark_user = g;
    return c;
}

// Device code indexes each array in 32 bits: the face index space (the
// largest array) must hold fewer than DG_MAX_ELEMS elements, which on one
// GPU's memory only the largest 1D grids exceed.
static bool fits_index_space(const dg_desc& d, int ghosts) {
    const long ex = d.nx + 1 + 2L * ghosts, ey = d.dim >= 2 ? d.ny + 1 + 2L * ghosts : 1, ez = d.dim >= 3 ? d.nz + 1 + 2L * ghosts : 1;
    if (ex >= DG_MAX_ELEMS || ey >= DG_MAX_ELEMS || ez >= DG_MAX_ELEMS) return false;
    const long exy = ex * ey;
    return exy < DG_MAX_ELEMS && exy * ez < DG_MAX_ELEMS;
}

extern "C" {

int dg_create(const dg_desc* desc, dg_grid** out) {
    dg_grid* g = nullptr;
    if (!desc || !out) return DG_ERR_ARG;
    const dg_desc& d = *desc;
    if (d.dim < 1 || d.dim > 3 || d.nx < 1 || (d.dim >= 2 && d.ny < 1) || (d.dim >= 3 && d.nz < 1)) return DG_ERR_ARG;
    if (!(d.dx > 0) || (d.dim >= 2 && !(d.dy > 0)) || (d.dim >= 3 && !(d.dz > 0))) return DG_ERR_ARG;
    if (d.ctu && !d.muscl) return DG_ERR_ARG;  // CTU.cpp:26-30
    if (!fits_index_space(d, std::max(d.ghosts, 3))) return DG_ERR_ARG;
    g = new dg_grid();
    g->d = d;
    if (g->d.timestep_tol <= 0) g->d.timestep_tol = 1e-14;
    // validGhosts (Splitting.cpp:19-27)
    int gh = d.ghosts;
    if (d.mhd && d.ctu) gh = std::max(gh, 3);
    else if (d.muscl) gh = std::max(gh, 2);
    else gh = std::max(gh, 1);
    g->g = gh;
    g->nf = d.mhd ? 8 : 5;
    g->ce = make_ext(d.dim, d.nx, d.ny, d.nz, gh);
    g->fe = make_ext(d.dim, d.nx + 1, d.ny + 1, d.nz + 1, gh);
    g->ph = make_phys(d);
    const long nc = g->ce.size, nfc = g->fe.size;
    const bool faces = d.mhd && d.dim >= 2;
    auto alloc = [&](double** p, long n) -> bool {
        if (n <= 0) return true;
        return hipMalloc((void**)p, sizeof(double) * n) == hipSuccess;
    };
    bool ok = hipStreamCreateWithFlags(&g->s, hipStreamNonBlocking) == hipSuccess;
    ok = ok && alloc(&g->W, 8 * nc) && alloc(&g->Wn, 8 * nc);
    // fused sweeps (MHD 2D/3D, hydro 2D) need predictor + corrector fluxes;
    // 1D and 3D hydro keep the reference's half-state arrays
    const bool fusable = (d.mhd && d.dim >= 2) || (!d.mhd && d.dim == 2);
    g->fused = fusable && getenv("DG_UNFUSED") == nullptr;
    // diagnostic: fixed faces per walker of the strided-axis sweeps (results do not depend on it)
    if (const char* mc = getenv("DG_MARCH_CHUNK")) g->march_chunk = atoi(mc);
    // check: the FP64 walkers take the register-window form the FP32 build
    // uses, so the FP64 parity tests hold that form to the oracle bit for bit
    g->march_window = getenv("DG_MARCH_WINDOW_CHECK") != nullptr;
    ok = ok && alloc(&g->F, (long)d.dim * g->nf * nc);
    if (g->fused) ok = ok && (!d.ctu || alloc(&g->Fp, (long)d.dim * g->nf * nc));
    else ok = ok && alloc(&g->HS, 2L * d.dim * g->nf * nc);
    if (!d.mhd && d.dim == 3 && d.ctu) ok = ok && alloc(&g->T, 5L * g->nf * nc);
    if (faces) ok = ok && alloc(&g->B, 3 * nfc) && alloc(&g->Bn, 3 * nfc) && alloc(&g->Bh, 3 * nfc) && alloc(&g->E0, 3 * nfc) &&
                    alloc(&g->Eh, 3 * nfc);
    if (d.dim >= 2) {
        g->ncflpart = ((long)nc + 255) / 256;
        ok = ok && alloc(&g->cflpart, g->ncflpart);
    }
    ok = ok && hipMalloc((void**)&g->dev_words, 4 * sizeof(unsigned long long)) == hipSuccess;
    ok = ok && hipHostMalloc((void**)&g->host_words, 4 * sizeof(unsigned long long), hipHostMallocDefault) == hipSuccess;
    if (!ok) {
        dg_destroy(g);
        return DG_ERR_HIP;
    }
    ok = hipMemsetAsync(g->W, 0, sizeof(double) * 8 * nc, g->s) == hipSuccess;
    ok = ok && hipMemsetAsync(g->Wn, 0, sizeof(double) * 8 * nc, g->s) == hipSuccess;
    if (faces) {
        ok = ok && hipMemsetAsync(g->B, 0, sizeof(double) * 3 * nfc, g->s) == hipSuccess;
        ok = ok && hipMemsetAsync(g->Bn, 0, sizeof(double) * 3 * nfc, g->s) == hipSuccess;
    }
    ok = ok && hipStreamSynchronize(g->s) == hipSuccess;
    if (!ok) {
        dg_destroy(g);
        return DG_ERR_HIP;
    }
    *out = g;
    return DG_OK;
}

int dg_destroy(dg_grid* g) {
    if (!g) return DG_OK;
    int rc = DG_OK;
    if (g->s && hipStreamSynchronize(g->s) != hipSuccess) rc = DG_ERR_HIP;
    if (g->cs && hipStreamSynchronize(g->cs) != hipSuccess) rc = DG_ERR_HIP;
    double* bufs[] = {g->W, g->Wn, g->B, g->Bn, g->Bh, g->HS, g->F, g->Fp, g->T, g->E0, g->Eh, g->cflpart, g->Ws, g->snap};
    for (double* p : bufs)
        if (p && hipFree(p) != hipSuccess) rc = DG_ERR_HIP;
    if (g->dev_words && hipFree(g->dev_words) != hipSuccess) rc = DG_ERR_HIP;
    if (g->host_words && hipHostFree(g->host_words) != hipSuccess) rc = DG_ERR_HIP;
    if (g->ev0 && hipEventDestroy(g->ev0) != hipSuccess) rc = DG_ERR_HIP;
    if (g->ev1 && hipEventDestroy(g->ev1) != hipSuccess) rc = DG_ERR_HIP;
    for (hipEvent_t e : g->evs)
        if (hipEventDestroy(e) != hipSuccess) rc = DG_ERR_HIP;
    if (g->snap_ready && hipEventDestroy(g->snap_ready) != hipSuccess) rc = DG_ERR_HIP;
    if (g->snap_done && hipEventDestroy(g->snap_done) != hipSuccess) rc = DG_ERR_HIP;
    if (g->cs && hipStreamDestroy(g->cs) != hipSuccess) rc = DG_ERR_HIP;
    if (g->s && hipStreamDestroy(g->s) != hipSuccess) rc = DG_ERR_HIP;
    delete g;
    return rc;
}

int dg_ghosts(const dg_grid* g) { return g ? g->g : -1; }
long dg_cell_count(const dg_grid* g) { return g ? g->ce.size : -1; }
long dg_face_count(const dg_grid* g) { return g && g->d.mhd && g->d.dim >= 2 ? g->fe.size : 0; }
const char* dg_last_error(const dg_grid* g) { return g ? g->err.c_str() : "null handle"; }
void* dg_stream(dg_grid* g) { return g ? (void*)g->s : nullptr; }

// A staging buffer that is freed on every return path.  release() frees it
// with the error check of the success path.
struct Staging {
    double* p = nullptr;
    ~Staging() {
        if (p) (void)hipFree(p);
    }
    hipError_t release() {
        double* q = p;
        p = nullptr;
        return q ? hipFree(q) : hipSuccess;
    }
};

int dg_upload(dg_grid* g, const double* w, int stride, const double* B) {
    if (!g || !w || stride < g->nf) return DG_ERR_ARG;
    g->cfl_cached = false;
    const long nc = g->ce.size;
    Staging stage;
    HIPCHK(hipMalloc((void**)&stage.p, sizeof(double) * nc * stride));
    HIPCHK(hipMemcpyAsync(stage.p, w, sizeof(double) * nc * stride, hipMemcpyHostToDevice, g->s));
    launch_aos_to_soa(stage.p, stride, 8, soa(g->W, nc), nc, g->s);
    if (B && g->B) {
        const long nfc = g->fe.size;
        Staging sb;
        HIPCHK(hipMalloc((void**)&sb.p, sizeof(double) * 3 * nfc));
        HIPCHK(hipMemcpyAsync(sb.p, B, sizeof(double) * 3 * nfc, hipMemcpyHostToDevice, g->s));
        launch_aos_to_soa(sb.p, 3, 3, soa(g->B, nfc), nfc, g->s);
        HIPCHK(hipStreamSynchronize(g->s));
        HIPCHK(sb.release());
    }
    HIPCHK(hipStreamSynchronize(g->s));
    HIPCHK(stage.release());
    return DG_OK;
}

int dg_download(dg_grid* g, double* w, int stride, double* B) {
    if (!g || !w || stride < g->nf) return DG_ERR_ARG;
    const long nc = g->ce.size;
    Staging stage;
    HIPCHK(hipMalloc((void**)&stage.p, sizeof(double) * nc * stride));
    launch_soa_to_aos(soa(g->W, nc), g->d.mhd ? 8 : 5, stage.p, stride, nc, g->s);
    HIPCHK(hipMemcpyAsync(w, stage.p, sizeof(double) * nc * stride, hipMemcpyDeviceToHost, g->s));
    HIPCHK(hipStreamSynchronize(g->s));
    HIPCHK(stage.release());
    if (B && g->B) {
        const long nfc = g->fe.size;
        Staging sb;
        HIPCHK(hipMalloc((void**)&sb.p, sizeof(double) * 3 * nfc));
        launch_soa_to_aos(soa(g->B, nfc), 3, sb.p, 3, nfc, g->s);
        HIPCHK(hipMemcpyAsync(B, sb.p, sizeof(double) * 3 * nfc, hipMemcpyDeviceToHost, g->s));
        HIPCHK(hipStreamSynchronize(g->s));
        HIPCHK(sb.release());
    }
    return DG_OK;
}

int dg_upload_device(dg_grid* g, const double* w_soa, const double* B_soa) {
    if (!g || !w_soa) return DG_ERR_ARG;
    g->cfl_cached = false;
    HIPCHK(hipMemcpyAsync(g->W, w_soa, sizeof(double) * 8 * g->ce.size, hipMemcpyDeviceToDevice, g->s));
    if (B_soa && g->B) HIPCHK(hipMemcpyAsync(g->B, B_soa, sizeof(double) * 3 * g->fe.size, hipMemcpyDeviceToDevice, g->s));
    HIPCHK(hipStreamSynchronize(g->s));
    return DG_OK;
}

int dg_download_device(dg_grid* g, double* w_soa, double* B_soa) {
    if (!g || !w_soa) return DG_ERR_ARG;
    HIPCHK(hipMemcpyAsync(w_soa, g->W, sizeof(double) * 8 * g->ce.size, hipMemcpyDeviceToDevice, g->s));
    if (B_soa && g->B) HIPCHK(hipMemcpyAsync(B_soa, g->B, sizeof(double) * 3 * g->fe.size, hipMemcpyDeviceToDevice, g->s));
    HIPCHK(hipStreamSynchronize(g->s));
    return DG_OK;
}

int dg_set_boundary(dg_grid* g, const dg_bc* list, int n) {
    if (!g || n < 0 || (n > 0 && !list)) return DG_ERR_ARG;
    g->bcs.clear();
    for (int q = 0; q < n; q++) {
        dg_bc b = list[q];
        if (b.kind < 0 || b.kind > DG_BC_IGNORE) return fail(g, DG_ERR_ARG, "unknown boundary kind");
        if (b.kind == DG_BC_PERIODIC)  // Periodic.cpp:14-15
            b.faces = ((b.faces & 3) ? 3 : 0) | ((b.faces & 12) ? 12 : 0) | ((b.faces & 48) ? 48 : 0);
        g->bcs.push_back(b);
    }
    return DG_OK;
}

// One element of the list, in the order of its reference apply(): cell passes
// by face, then (MHD 2D/3D) face-B passes by face.
static void apply_one(dg_grid* g, const dg_bc& b) {
    if (b.kind == DG_BC_IGNORE) return;
    const int D = g->d.dim, ng = g->g;
    const bool faces = g->d.mhd && D >= 2;
    BCArgs a{};
    a.W = soa(g->W, g->ce.size);
    for (int q = 0; q < 3; q++) a.B[q] = faces ? g->B + q * g->fe.size : nullptr;
    a.ce = g->ce; a.fe = g->fe;
    a.dim = D; a.nf = g->d.mhd ? 8 : 5; a.kind = b.kind; a.corners = b.corners; a.gated = b.gated;
    for (int q = 0; q < 8; q++) a.st[q] = b.state[q];
    const int n[3] = {g->ce.n[0], g->ce.n[1], g->ce.n[2]};
    auto range = [&](int d, bool facepass, int& lo, int& hi) {
        if (d >= D) { lo = 0; hi = 1; return; }
        lo = b.corners ? -ng : 0;
        hi = (b.corners ? n[d] + ng : n[d]) + (facepass ? 1 : 0);
    };
    for (int pass = 0; pass < (faces ? 2 : 1); pass++) {
        const bool fp = pass == 1;
        for (int ax = 0; ax < D; ax++) {
            int t[2], c = 0;
            for (int d = 0; d < 3; d++)
                if (d != ax) t[c++] = d;
            a.axis = ax;
            range(t[0], fp, a.lo[0], a.hi[0]);
            range(t[1], fp, a.lo[1], a.hi[1]);
            const int neg = 1 << (2 * ax), pos = 2 << (2 * ax);
            if (b.kind == DG_BC_PERIODIC) {
                if (b.faces & (neg | pos)) { a.side = 2; launch_bc(a, fp, g->s); }
                continue;
            }
            if (b.faces & neg) { a.side = 0; launch_bc(a, fp, g->s); }
            if (b.faces & pos) { a.side = 1; launch_bc(a, fp, g->s); }
        }
    }
}

// One element applied alone, without the list's implicit Outflow: the step
// by which a caller interleaves device elements with elements it evaluates
// itself on the host (user GhostFill subclasses).
int dg_apply_boundary_element(dg_grid* g, const dg_bc* bc) {
    if (!g || !bc) return DG_ERR_ARG;
    dg_bc b = *bc;
    if (b.kind < 0 || b.kind > DG_BC_IGNORE) return fail(g, DG_ERR_ARG, "unknown boundary kind");
    if (b.kind == DG_BC_PERIODIC) b.faces = ((b.faces & 3) ? 3 : 0) | ((b.faces & 12) ? 12 : 0) | ((b.faces & 48) ? 48 : 0);
    apply_one(g, b);  // ghosts only: the apply pass's CFL maxima stay valid, as for dg_apply_boundary
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(g, DG_ERR_HIP, hipGetErrorString(e));
    return DG_OK;
}

int dg_set_boundary_callback(dg_grid* g, int (*cb)(void* user), void* user) {
    if (!g) return DG_ERR_ARG;
    g->bc_cb = cb;
    g->bc_user = user;
    return DG_OK;
}

int dg_apply_boundary(dg_grid* g) {
    if (!g) return DG_ERR_ARG;
    if (g->bc_cb) {  // the caller applies the list itself (host-evaluated elements)
        g->cfl_cached = false;
        const int rc = g->bc_cb(g->bc_user);
        return rc ? fail(g, rc, "boundary callback failed") : DG_OK;
    }
    int uncovered = 63;  // BoundaryList.hpp:104-110
    for (auto& b : g->bcs) uncovered &= ~b.faces;
    dg_bc imp{};
    imp.kind = DG_BC_OUTFLOW; imp.faces = uncovered; imp.corners = 1; imp.gated = 0;
    tm_record(g, -1);
    apply_one(g, imp);
    for (auto& b : g->bcs) apply_one(g, b);
    tm_record(g, K_BC);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(g, DG_ERR_HIP, hipGetErrorString(e));
    return DG_OK;
}

int dg_cfl_dt(dg_grid* g, double* dt, dg_status* st) {
    if (!g || !dt) return DG_ERR_ARG;
    launch_fill_u64(g->dev_words + 1, 0ull, 1, g->s);
    StepCtx c = make_ctx(g, 0.0);
    tm_record(g, -1);
    if (g->cfl_cached) launch_cfl_tail(c, g->ncflpart, g->dev_words + 1, g->s);
    else launch_cfl(c, g->dev_words + 1, g->s);
    tm_record(g, K_CFL);
    HIPCHK(hipMemcpyAsync(g->host_words + 1, g->dev_words + 1, sizeof(unsigned long long), hipMemcpyDeviceToHost, g->s));
    HIPCHK(hipStreamSynchronize(g->s));
    tm_flush(g);
    unsigned long long bits = g->host_words[1];
    double m;
    memcpy(&m, &bits, sizeof m);
    if (!(m > 0)) {
        const char* msg = "CFL timestep failed: max signal speed is zero";
        set_status(st, DG_ERR_CFL, 0, 0, 0, msg);
        return fail(g, DG_ERR_CFL, msg);
    }
    *dt = g->d.dim == 1 ? g->d.cfl_coeff * g->d.dx / m : g->d.cfl_coeff / m;
    set_status(st, DG_OK, 0, 0, 0, "");
    return DG_OK;
}

static void decode(const dg_grid* g, unsigned long long key, dg_status* st, int* code, std::string* msg) {
    const int stage = (int)(key >> 56), ec = (int)(key & 0xff);
    const long m = (long)((key >> 8) & 0xFFFFFFFFFFFFull);
    const Ext& e = g->ce;
    int ijk[3] = {0, 0, 0};
    long r = m;
    const int D = g->d.dim;
    const int Ny = D >= 2 ? e.n[1] + 2 * e.g : 1, Nz = D >= 3 ? e.n[2] + 2 * e.g : 1;
    if (D == 3) { ijk[2] = (int)(r % Nz) - e.g; r /= Nz; }
    if (D >= 2) { ijk[1] = (int)(r % Ny) - e.g; r /= Ny; }
    ijk[0] = (int)r - e.g;
    char b[192];
    auto where = [&]() {
        if (D == 1) snprintf(b, sizeof b, "(%d)", ijk[0]);
        else if (D == 2) snprintf(b, sizeof b, "(%d,%d)", ijk[0], ijk[1]);
        else snprintf(b, sizeof b, "(%d,%d,%d)", ijk[0], ijk[1], ijk[2]);
        return std::string(b);
    };
    (void)stage;
    switch (ec) {
        case E_RIEMANN_INPUT: *code = DG_ERR_RIEMANN_INPUT; *msg = "Riemann Solution is unphysical because inputs were not physical"; break;
        case E_RIEMANN_SOLUTION: *code = DG_ERR_RIEMANN_SOLUTION; *msg = "Riemann Solution to physical inputs is unphysical"; break;
        case E_NAN: *code = DG_ERR_NAN; *msg = "\tNaN state would be produced at " + where() + "\n"; break;
        default: *code = DG_ERR_UNPHYSICAL; *msg = "Unphysical state would be produced at " + where(); break;
    }
    set_status(st, *code, ijk[0], ijk[1], ijk[2], *msg);
}

int dg_step(dg_grid* g, double dt, dg_status* st) {
    if (!g) return DG_ERR_ARG;
    g->cfl_cached = false;  // the apply pass overwrites the maxima
    launch_fill_u64(g->dev_words, NO_ERR, 1, g->s);
    StepCtx c = make_ctx(g, dt);
    if (g->timing) HIPCHK(hipEventRecord(g->ev0, g->s));
    tm_record(g, -1);
    if (c.mixed) p32::launch_step(c, g->s);
    else p64::launch_step(c, g->s);
    HIPCHK(hipGetLastError());
    if (g->timing) HIPCHK(hipEventRecord(g->ev1, g->s));
    HIPCHK(hipMemcpyAsync(g->host_words, g->dev_words, sizeof(unsigned long long), hipMemcpyDeviceToHost, g->s));
    HIPCHK(hipStreamSynchronize(g->s));
    if (g->timing) {
        float ms = 0;
        HIPCHK(hipEventElapsedTime(&ms, g->ev0, g->ev1));
        g->step_ms += ms;
        g->timed_steps++;
        tm_flush(g);
    }
    const unsigned long long key = g->host_words[0];
    if (key != NO_ERR) {
        int code;
        std::string msg;
        decode(g, key, st, &code, &msg);
        return fail(g, code, msg);
    }
    std::swap(g->W, g->Wn);
    if (g->B) std::swap(g->B, g->Bn);
    g->cfl_cached = g->cflpart != nullptr;
    set_status(st, DG_OK, 0, 0, 0, "");
    return DG_OK;
}

// Grid2D/3D::split_step (Splitting.cpp:118-216): Strang splitting of hydro,
// the sweep order rotating with sweep_step (2 orders in 2D, 6 in 3D).  The
// sweeps run W -> Wn -> Ws -> Wn [-> Ws -> Wn], each after the boundary pass
// on its input; W is kept, so a failed step commits nothing.
int dg_split_step(dg_grid* g, double dt, dg_status* st) {
    if (!g) return DG_ERR_ARG;
    const int D = g->d.dim;
    if (D == 1) return dg_step(g, dt, st);  // Grid1D::split_step is unsplit_step (Splitting.cpp:98)
    if (g->d.mhd) {
        const char* msg = "MHD requires Unsplit dimensions. Please enable DIMENSION_UNSPLIT in Config.h";  // Electric.cpp:17-18
        set_status(st, DG_ERR_UNSUPPORTED, 0, 0, 0, msg);
        return fail(g, DG_ERR_UNSUPPORTED, msg);
    }
    // 8 fields like W and Wn: a boundary callback between sweeps sees Ws as
    // the grid state, and dg_upload / dg_download move 8 fields
    if (!g->Ws && hipMalloc((void**)&g->Ws, sizeof(double) * 8 * g->ce.size) != hipSuccess) return fail(g, DG_ERR_HIP, "hipMalloc");
    g->cfl_cached = false;
    launch_fill_u64(g->dev_words, NO_ERR, 1, g->s);
    static const int order3[6][3] = {{0, 1, 2}, {2, 0, 1}, {1, 2, 0}, {2, 1, 0}, {0, 2, 1}, {1, 0, 2}};  // Splitting.cpp:160-201
    int ax[5], n = 0;
    double h[5];
    if (D == 2) {
        const int a = g->sweep_step % 2 == 0 ? 0 : 1, b = 1 - a;
        ax[0] = a; ax[1] = b; ax[2] = a;
        h[0] = dt / 2; h[1] = dt; h[2] = dt / 2;
        n = 3;
    } else {
        const int* o = order3[g->sweep_step % 6];
        ax[0] = o[0]; ax[1] = o[1]; ax[2] = o[2]; ax[3] = o[1]; ax[4] = o[0];
        h[0] = dt / 2; h[1] = dt / 2; h[2] = dt; h[3] = dt / 2; h[4] = dt / 2;
        n = 5;
    }
    if (g->timing) HIPCHK(hipEventRecord(g->ev0, g->s));
    double* in = g->W;
    double* const W0 = g->W;
    for (int q = 0; q < n; q++) {
        double* out = q % 2 == 0 ? g->Wn : g->Ws;
        g->W = in;  // the boundary pass runs on this sweep's input
        int rc = dg_apply_boundary(g);
        g->W = W0;
        if (rc) return rc;
        StepCtx c = make_ctx(g, h[q]);
        tm_record(g, -1);
        if (c.mixed) p32::launch_split_axis(c, ax[q], soa(in, g->ce.size), soa(out, g->ce.size), g->s);
        else p64::launch_split_axis(c, ax[q], soa(in, g->ce.size), soa(out, g->ce.size), g->s);
        tm_record(g, K_FLUX_X + ax[q]);
        in = out;
    }
    HIPCHK(hipGetLastError());
    if (g->timing) HIPCHK(hipEventRecord(g->ev1, g->s));
    HIPCHK(hipMemcpyAsync(g->host_words, g->dev_words, sizeof(unsigned long long), hipMemcpyDeviceToHost, g->s));
    HIPCHK(hipStreamSynchronize(g->s));
    if (g->timing) {
        float ms = 0;
        HIPCHK(hipEventElapsedTime(&ms, g->ev0, g->ev1));
        g->step_ms += ms;
        g->timed_steps++;
        tm_flush(g);
    }
    const unsigned long long key = g->host_words[0];
    if (key != NO_ERR) {
        int code;
        std::string msg;
        decode(g, key, st, &code, &msg);
        return fail(g, code, msg);
    }
    std::swap(g->W, g->Wn);  // the last sweep wrote Wn (an odd number of sweeps)
    g->sweep_step++;
    set_status(st, DG_OK, 0, 0, 0, "");
    return DG_OK;
}

// Grid::advance_unsplit / advance_split (CFL.cpp:160-238)
static int advance_loop(dg_grid* g, double dt, int check_cfl, long max_steps, dg_stats* stats, dg_status* st, bool split) {
    dg_stats local{};
    dg_stats& S = stats ? *stats : local;
    S.steps = 0; S.failures = 0; S.t_advanced = 0;
    const double tol = g->d.timestep_tol;
    while (dt > tol) {
        if (max_steps >= 0 && S.steps >= max_steps) break;
        int rc = dg_apply_boundary(g);
        if (rc) return rc;
        double t1 = dt;
        if (check_cfl) {
            double c;
            rc = dg_cfl_dt(g, &c, st);
            if (rc) return rc;
            t1 = (c < dt) ? c : dt;  // std::min(dt, cfl)
        }
        while (true) {
            rc = split ? dg_split_step(g, t1, st) : dg_step(g, t1, st);
            if (rc == DG_OK) break;
            if (rc == DG_ERR_HIP || rc == DG_ERR_ARG || rc == DG_ERR_UNSUPPORTED) return rc;
            S.failures++;
            t1 *= 0.5;
            if (t1 < tol) {
                const char* msg = "Timestep has fallen below minimum. Exiting";
                set_status(st, DG_ERR_DT_UNDERFLOW, 0, 0, 0, msg);
                return fail(g, DG_ERR_DT_UNDERFLOW, msg);
            }
        }
        dt -= t1;
        S.t_advanced += t1;
        S.steps++;
    }
    set_status(st, DG_OK, 0, 0, 0, "");
    return DG_OK;
}

int dg_advance(dg_grid* g, double dt, int check_cfl, long max_steps, dg_stats* stats, dg_status* st) {
    if (!g) return DG_ERR_ARG;
    return advance_loop(g, dt, check_cfl, max_steps, stats, st, false);
}

int dg_advance_split(dg_grid* g, double dt, int check_cfl, long max_steps, dg_stats* stats, dg_status* st) {
    if (!g) return DG_ERR_ARG;
    return advance_loop(g, dt, check_cfl, max_steps, stats, st, true);
}

int dg_init_B(dg_grid* g) {
    if (!g) return DG_ERR_ARG;
    g->cfl_cached = false;
    int rc = dg_apply_boundary(g);
    if (rc) return rc;
    if (!(g->d.mhd && g->d.dim >= 2)) return DG_OK;
    launch_init_B(g->d.dim, soa(g->W, g->ce.size), g->B, g->ce, g->fe, g->s);
    HIPCHK(hipStreamSynchronize(g->s));
    return DG_OK;
}

int dg_timing(dg_grid* g, int enable, double* step_ms, long* steps) {
    if (!g) return DG_ERR_ARG;
    if (step_ms) *step_ms = g->step_ms;
    if (steps) *steps = g->timed_steps;
    if (enable >= 0) {
        if (enable && !g->ev0) {
            HIPCHK(hipEventCreate(&g->ev0));
            HIPCHK(hipEventCreate(&g->ev1));
        }
        HIPCHK(hipStreamSynchronize(g->s));
        tm_flush(g);
        g->timing = enable != 0;
        g->step_ms = 0;
        g->timed_steps = 0;
        for (int k = 0; k < K_COUNT; k++) g->kms[k] = 0, g->klaunch[k] = 0;
    }
    return DG_OK;
}

int dg_kernel_time(dg_grid* g, int id, const char** name, double* ms, long* launches) {
    if (!g || id < 0 || id >= K_COUNT) return DG_ERR_ARG;
    if (name) *name = kKernelNames[id];
    if (ms) *ms = g->kms[id];
    if (launches) *launches = g->klaunch[id];
    return DG_OK;
}

int dg_divb(dg_grid* g, double* out) {
    if (!g || !out) return DG_ERR_ARG;
    *out = 0;
    if (!(g->d.mhd && g->d.dim >= 2)) return DG_OK;
    launch_fill_u64(g->dev_words + 2, 0ull, 2, g->s);
    const double d[3] = {g->d.dx, g->d.dy, g->d.dz};
    launch_divb(g->d.dim, g->B, g->ce, g->fe, d, g->dev_words + 2, g->s);
    HIPCHK(hipMemcpyAsync(g->host_words + 2, g->dev_words + 2, 2 * sizeof(unsigned long long), hipMemcpyDeviceToHost, g->s));
    HIPCHK(hipStreamSynchronize(g->s));
    double dv, bm;
    memcpy(&dv, &g->host_words[2], 8);
    memcpy(&bm, &g->host_words[3], 8);
    double dmin = g->d.dx;
    if (g->d.dim >= 2) dmin = std::min(dmin, g->d.dy);
    if (g->d.dim >= 3) dmin = std::min(dmin, g->d.dz);
    *out = bm > 0 ? dv * dmin / bm : 0.0;
    return DG_OK;
}

}  // extern "C"

// ================================================ frame snapshot / restore
// The frame files store each field x fastest (HDF5Output.cpp:305-342); the
// device arrays are last-index fastest.  One field of a box of A x Bm x C
// points (a: x, b: y in 3D, c: the contiguous axis) is transposed through a
// 64 x 64 tile in LDS, so the reads and the writes are both coalesced.
struct XfastView {
    long sa, sb, sc, off;  // device strides of a, b, c and the offset of the box origin
    long A, Bm, C;         // box extents
};
__global__ __launch_bounds__(256) void k_to_xfast(const double* src, XfastView v, double* dst) {
    __shared__ double tile[64][65];
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
    const long a0 = (long)blockIdx.x * 64, c0 = (long)blockIdx.y * 64, b = blockIdx.z;
    for (int r = ty; r < 64; r += 4) {
        const long a = a0 + r, c = c0 + tx;
        if (a < v.A && c < v.C) tile[r][tx] = src[v.off + a * v.sa + b * v.sb + c * v.sc];
    }
    __syncthreads();
    for (int r = ty; r < 64; r += 4) {
        const long c = c0 + r, a = a0 + tx;
        if (a < v.A && c < v.C) dst[(c * v.Bm + b) * v.A + a] = tile[tx][r];
    }
}
__global__ __launch_bounds__(256) void k_from_xfast(const double* src, XfastView v, double* dst) {
    __shared__ double tile[64][65];
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
    const long a0 = (long)blockIdx.x * 64, c0 = (long)blockIdx.y * 64, b = blockIdx.z;
    for (int r = ty; r < 64; r += 4) {
        const long c = c0 + r, a = a0 + tx;
        if (a < v.A && c < v.C) tile[tx][r] = src[(c * v.Bm + b) * v.A + a];
    }
    __syncthreads();
    for (int r = ty; r < 64; r += 4) {
        const long a = a0 + r, c = c0 + tx;
        if (a < v.A && c < v.C) dst[v.off + a * v.sa + b * v.sb + c * v.sc] = tile[r][tx];
    }
}
// The box [-ng, n + ng) (cells) of an Ext, dims per the grid's dimension.
static XfastView xfast_view(const Ext& e, int dim, int ng) {
    XfastView v{};
    v.off = e.at(-ng, dim >= 2 ? -ng : 0, dim >= 3 ? -ng : 0);
    if (dim == 3) {
        v.sa = e.s[0]; v.sb = e.s[1]; v.sc = e.s[2];
        v.A = e.n[0] + 2 * ng; v.Bm = e.n[1] + 2 * ng; v.C = e.n[2] + 2 * ng;
    } else if (dim == 2) {
        v.sa = e.s[0]; v.sb = 0; v.sc = e.s[1];
        v.A = e.n[0] + 2 * ng; v.Bm = 1; v.C = e.n[1] + 2 * ng;
    } else {
        v.sa = e.s[0]; v.sb = 0; v.sc = 0;
        v.A = e.n[0] + 2 * ng; v.Bm = 1; v.C = 1;
    }
    return v;
}
static void launch_xfast(bool to, const double* src, const XfastView& v, double* dst, hipStream_t s) {
    const dim3 grid((unsigned)((v.A + 63) / 64), (unsigned)((v.C + 63) / 64), (unsigned)v.Bm);
    if (to) hipLaunchKernelGGL(k_to_xfast, grid, dim3(256), 0, s, src, v, dst);
    else hipLaunchKernelGGL(k_from_xfast, grid, dim3(256), 0, s, src, v, dst);
}

extern "C" {

int dg_snapshot(dg_grid* g, int ng, double* cells, double* faces) {
    if (!g || !cells || ng < 0 || ng > g->g) return DG_ERR_ARG;
    const int D = g->d.dim;
    const bool fb = g->d.mhd && D >= 2;
    const XfastView vc = xfast_view(g->ce, D, ng), vf = xfast_view(g->fe, D, ng);
    const long N = vc.A * vc.Bm * vc.C, NF = fb ? vf.A * vf.Bm * vf.C : 0;
    const long need = 8 * N + 3 * NF;
    if (!g->cs) {
        HIPCHK(hipStreamCreateWithFlags(&g->cs, hipStreamNonBlocking));
        HIPCHK(hipEventCreateWithFlags(&g->snap_ready, hipEventDisableTiming));
        HIPCHK(hipEventCreateWithFlags(&g->snap_done, hipEventDisableTiming));
    }
    if (g->snap_recorded) HIPCHK(hipStreamWaitEvent(g->s, g->snap_done, 0));  // staging may still be copied out
    if (g->snap_len < need) {
        HIPCHK(hipStreamSynchronize(g->s));
        if (g->snap) HIPCHK(hipFree(g->snap));
        g->snap = nullptr;
        g->snap_len = 0;
        HIPCHK(hipMalloc((void**)&g->snap, sizeof(double) * need));
        g->snap_len = need;
    }
    const long nc = g->ce.size, nfc = g->fe.size;
    for (int f = 0; f < 8; f++) launch_xfast(true, g->W + f * nc, vc, g->snap + f * N, g->s);
    if (fb)
        for (int c = 0; c < 3; c++) launch_xfast(true, g->B + c * nfc, vf, g->snap + 8 * N + c * NF, g->s);
    HIPCHK(hipGetLastError());
    HIPCHK(hipEventRecord(g->snap_ready, g->s));
    HIPCHK(hipStreamWaitEvent(g->cs, g->snap_ready, 0));
    HIPCHK(hipMemcpyAsync(cells, g->snap, sizeof(double) * 8 * N, hipMemcpyDeviceToHost, g->cs));
    if (fb && faces) HIPCHK(hipMemcpyAsync(faces, g->snap + 8 * N, sizeof(double) * 3 * NF, hipMemcpyDeviceToHost, g->cs));
    HIPCHK(hipEventRecord(g->snap_done, g->cs));
    g->snap_recorded = true;
    return DG_OK;
}

// Touches no handle state but the event, so it may run on another thread
// while this one keeps stepping the grid.
int dg_snapshot_wait(dg_grid* g) {
    if (!g) return DG_ERR_ARG;
    if (g->snap_done) HIPCHK(hipEventSynchronize(g->snap_done));
    return DG_OK;
}

int dg_restore(dg_grid* g, int ng, const double* cells, const double* faces) {
    if (!g || !cells || ng < 0 || ng > g->g) return DG_ERR_ARG;
    const int D = g->d.dim;
    const bool fb = g->d.mhd && D >= 2 && faces;
    const XfastView vc = xfast_view(g->ce, D, ng), vf = xfast_view(g->fe, D, ng);
    const long N = vc.A * vc.Bm * vc.C, NF = fb ? vf.A * vf.Bm * vf.C : 0;
    if (int rc = dg_snapshot_wait(g)) return rc;
    g->cfl_cached = false;
    Staging stage;
    HIPCHK(hipMalloc((void**)&stage.p, sizeof(double) * (8 * N + 3 * NF)));
    HIPCHK(hipMemcpyAsync(stage.p, cells, sizeof(double) * 8 * N, hipMemcpyHostToDevice, g->s));
    if (fb) HIPCHK(hipMemcpyAsync(stage.p + 8 * N, faces, sizeof(double) * 3 * NF, hipMemcpyHostToDevice, g->s));
    const long nc = g->ce.size, nfc = g->fe.size;
    for (int f = 0; f < g->nf; f++) launch_xfast(false, stage.p + f * N, vc, g->W + f * nc, g->s);
    if (fb)
        for (int c = 0; c < 3; c++) launch_xfast(false, stage.p + 8 * N + c * NF, vf, g->B + c * nfc, g->s);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(g->s));
    HIPCHK(stage.release());
    return DG_OK;
}

int dg_set_spacing(dg_grid* g, double dx, double dy, double dz) {
    if (!g || !(dx > 0) || (g->d.dim >= 2 && !(dy > 0)) || (g->d.dim >= 3 && !(dz > 0))) return DG_ERR_ARG;
    g->d.dx = dx;
    if (g->d.dim >= 2) g->d.dy = dy;
    if (g->d.dim >= 3) g->d.dz = dz;
    g->cfl_cached = false;
    return DG_OK;
}

// ====================================================== no-allocation step
// dg_step_noalloc: the unsplit step on caller-owned device buffers, with its
// scratch (fluxes, half-step faces, EMFs, the failure word) in this
// file-scope device array instead of allocations.  Each GPU holds its own
// copy; its address is looked up once per device (hipGetSymbolAddress) and
// only that address is handed to kernels and copies.
static constexpr long NOALLOC_SCRATCH_BYTES = 64L << 20;
static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];
static constexpr int NOALLOC_MAX_DEVICES = 64;
static std::atomic<void*> g_scratch_ptr[NOALLOC_MAX_DEVICES] = {};
static thread_local std::string g_noalloc_err;

// Scratch layout of one step (the arrays dg_create allocates for the same
// grid, minus the state): doubles from the start of the array.
struct NoallocLayout {
    long words = 0, F = 0, Fp = -1, HS = -1, T = -1, Bh = -1, E0 = -1, Eh = -1, end = 0;
};
static NoallocLayout noalloc_layout(const dg_desc& d, const Ext& ce, const Ext& fe, bool fused) {
    NoallocLayout L;
    const long nc = ce.size, nfc = fe.size, nf = d.mhd ? 8 : 5;
    long at = 4;  // failure key + 3 spare words (u64, the size of a double)
    auto take = [&](long n) { const long o = at; at += (n + 31) / 32 * 32; return o; };
    L.words = 0;
    L.F = take((long)d.dim * nf * nc);
    if (fused) { if (d.ctu) L.Fp = take((long)d.dim * nf * nc); }
    else L.HS = take(2L * d.dim * nf * nc);
    if (!d.mhd && d.dim == 3 && d.ctu) L.T = take(5L * nf * nc);
    if (d.mhd && d.dim >= 2) { L.Bh = take(3 * nfc); L.E0 = take(3 * nfc); L.Eh = take(3 * nfc); }
    L.end = at;
    return L;
}

extern "C" long dg_noalloc_scratch_bytes(const dg_desc* desc) {
    if (!desc || desc->dim < 1 || desc->dim > 3) return -1;
    dg_desc d = *desc;
    int gh = d.ghosts;
    if (d.mhd && d.ctu) gh = std::max(gh, 3);
    else if (d.muscl) gh = std::max(gh, 2);
    else gh = std::max(gh, 1);
    const Ext ce = make_ext(d.dim, d.nx, d.ny, d.nz, gh), fe = make_ext(d.dim, d.nx + 1, d.ny + 1, d.nz + 1, gh);
    const bool fused = (d.mhd && d.dim >= 2) || (!d.mhd && d.dim == 2);
    return noalloc_layout(d, ce, fe, fused).end * (long)sizeof(double);
}
extern "C" long dg_noalloc_scratch_limit(void) { return NOALLOC_SCRATCH_BYTES; }
extern "C" const char* dg_noalloc_last_error(void) { return g_noalloc_err.c_str(); }

static int noalloc_fail(int code, const std::string& m) {
    g_noalloc_err = m;
    return code;
}

extern "C" int dg_step_noalloc(const dg_desc* desc, const double* w_in, const double* B_in, double* w_out, double* B_out,
                               double dt, void* stream, dg_status* st) {
    if (!desc || !w_in || !w_out) return noalloc_fail(DG_ERR_ARG, "dg_step_noalloc: null argument");
    const dg_desc& d = *desc;
    if (d.dim < 1 || d.dim > 3 || d.nx < 1 || (d.dim >= 2 && d.ny < 1) || (d.dim >= 3 && d.nz < 1))
        return noalloc_fail(DG_ERR_ARG, "dg_step_noalloc: bad sizes");
    if (!(d.dx > 0) || (d.dim >= 2 && !(d.dy > 0)) || (d.dim >= 3 && !(d.dz > 0)) || (d.ctu && !d.muscl))
        return noalloc_fail(DG_ERR_ARG, "dg_step_noalloc: bad spacing or switches");
    if (!fits_index_space(d, std::max(d.ghosts, 3))) return noalloc_fail(DG_ERR_ARG, "dg_step_noalloc: grid too large");
    const bool faces = d.mhd && d.dim >= 2;
    if (faces && (!B_in || !B_out)) return noalloc_fail(DG_ERR_ARG, "dg_step_noalloc: face B buffers required for 2D/3D MHD");
    if (w_in == w_out || (faces && B_in == B_out)) return noalloc_fail(DG_ERR_ARG, "dg_step_noalloc: outputs alias the inputs");
    const long need = dg_noalloc_scratch_bytes(desc);
    if (need > NOALLOC_SCRATCH_BYTES)
        return noalloc_fail(DG_ERR_ARG, "dg_step_noalloc: the grid needs " + std::to_string(need) + " bytes of scratch, more than the " +
                                            std::to_string(NOALLOC_SCRATCH_BYTES) + "-byte static array");
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= NOALLOC_MAX_DEVICES)
        return noalloc_fail(DG_ERR_HIP, "dg_step_noalloc: no current device");
    void* sp = g_scratch_ptr[dev].load(std::memory_order_acquire);
    if (!sp) {  // once per device, with that device current (a racing first call stores the same address)
        const hipError_t e = hipGetSymbolAddress(&sp, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !sp) return noalloc_fail(DG_ERR_HIP, std::string("hipGetSymbolAddress: ") + hipGetErrorString(e));
        g_scratch_ptr[dev].store(sp, std::memory_order_release);
    }
    double* base = (double*)sp;
    // a dg_grid on the stack describing the caller's buffers; nothing allocated
    dg_grid g;
    g.d = d;
    if (g.d.timestep_tol <= 0) g.d.timestep_tol = 1e-14;
    int gh = d.ghosts;
    if (d.mhd && d.ctu) gh = std::max(gh, 3);
    else if (d.muscl) gh = std::max(gh, 2);
    else gh = std::max(gh, 1);
    g.g = gh;
    g.nf = d.mhd ? 8 : 5;
    g.ce = make_ext(d.dim, d.nx, d.ny, d.nz, gh);
    g.fe = make_ext(d.dim, d.nx + 1, d.ny + 1, d.nz + 1, gh);
    g.ph = make_phys(d);
    g.fused = (d.mhd && d.dim >= 2) || (!d.mhd && d.dim == 2);
    const NoallocLayout L = noalloc_layout(d, g.ce, g.fe, g.fused);
    g.W = const_cast<double*>(w_in);
    g.Wn = w_out;
    if (faces) { g.B = const_cast<double*>(B_in); g.Bn = B_out; }
    g.F = base + L.F;
    if (L.Fp >= 0) g.Fp = base + L.Fp;
    if (L.HS >= 0) g.HS = base + L.HS;
    if (L.T >= 0) g.T = base + L.T;
    if (L.Bh >= 0) { g.Bh = base + L.Bh; g.E0 = base + L.E0; g.Eh = base + L.Eh; }
    g.dev_words = (unsigned long long*)(base + L.words);
    g.cflpart = nullptr;
    g.s = (hipStream_t)stream;
    // the failure word is the only scratch that must start at a set value
    launch_fill_u64(g.dev_words, NO_ERR, 1, g.s);
    StepCtx c = make_ctx(&g, dt);
    if (c.mixed) p32::launch_step(c, g.s);
    else p64::launch_step(c, g.s);
    hipError_t e = hipGetLastError();
    unsigned long long key = NO_ERR;
    if (e == hipSuccess) e = hipMemcpyAsync(&key, g.dev_words, sizeof key, hipMemcpyDeviceToHost, g.s);
    if (e == hipSuccess) e = hipStreamSynchronize(g.s);
    if (e != hipSuccess) return noalloc_fail(DG_ERR_HIP, std::string("dg_step_noalloc: ") + hipGetErrorString(e));
    if (key != NO_ERR) {
        int code;
        std::string msg;
        decode(&g, key, st, &code, &msg);
        return noalloc_fail(code, msg);
    }
    set_status(st, DG_OK, 0, 0, 0, "");
    return DG_OK;
}

int dg_host_alloc(long bytes, void** p) {
    if (!p || bytes <= 0) return DG_ERR_ARG;
    return hipHostMalloc(p, (size_t)bytes, hipHostMallocDefault) == hipSuccess ? DG_OK : DG_ERR_HIP;
}
int dg_host_free(void* p) { return !p || hipHostFree(p) == hipSuccess ? DG_OK : DG_ERR_HIP; }

}  // extern "C"

// ======================================================= batched patches
// Thousands of independent 2D MHD patches (CTU + CT), each with its own CFL
// dt, advanced together: every stage of the step is one launch over all
// patches (block tables map workgroups to patches), so one step is a fixed
// sequence of launches that a HIP graph replays.  Patch p is committed by
// flipping its parity word.
struct dg_batch {
    dg_desc d{};
    Phys ph{};
    int npatch = 0, bc_kind = 0;
    std::vector<PatchDesc> P;
    std::vector<int> order;                         // largest first: block tables list big patches first
    PatchDesc* dP = nullptr;
    double *W[2] = {nullptr, nullptr}, *B[2] = {nullptr, nullptr};
    double *Fp = nullptr, *F = nullptr, *Bh = nullptr, *E0 = nullptr, *Eh = nullptr;
    int* parity = nullptr;
    double* dt_last = nullptr;
    long *steps = nullptr, *fails = nullptr;
    int* status = nullptr;
    unsigned long long *err = nullptr, *vmax = nullptr, *vnext = nullptr;
    int* cfl_ok = nullptr;
    double* dtp = nullptr;
    int *halv = nullptr, *active = nullptr;
    double* stage = nullptr;                        // upload/download staging, largest patch
    int* bt_mem[BK_COUNT] = {};                     // per kind: [patch of each block | first item of each block]
    BlockTable bt[BK_COUNT] = {};
    long cells = 0, faces = 0, zones = 0, max_nc = 0, max_nfc = 0;
    int chunk[BATCH_CLASSES] = {8, 16, 32, 32};
    int ysw[BATCH_CLASSES] = {64, 64, 64, 64};   // y-sweep segment width per class
    hipStream_t s = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    int graph_steps = 0;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    std::string err_msg;
};

#define BCHK(x)                                                                \
    do {                                                                       \
        hipError_t e_ = (x);                                                   \
        if (e_ != hipSuccess) {                                                \
            if (b) b->err_msg = std::string(#x) + ": " + hipGetErrorString(e_); \
            return DG_ERR_HIP;                                                 \
        }                                                                      \
    } while (0)

static int batch_class(int nx, int ny) {
    const int n = std::max(nx, ny);
    return n < 64 ? 0 : n < 128 ? 1 : n < 192 ? 2 : 3;
}

// Work items of one patch for each kind (the ranges of the batched kernels).
static long batch_items(const dg_batch* b, const PatchDesc& q, int kind) {
    const long nx = q.n[0], ny = q.n[1], sides = b->bc_kind == DG_BC_PERIODIC ? 1 : 2;
    auto chunked = [&](long len, long rows) {
        const long c = std::min<long>(b->chunk[q.cls], len);
        return ((len + c - 1) / c) * rows;
    };
    auto segments = [&](long len, long rows) {  // whole wavefronts of SW-lane segments, or the cells streamed (SW 0)
        const long sw = b->ysw[q.cls];
        if (sw == 0) {
            const long cells = rows * (len + 1);
            return (cells > 1 ? (cells - 1 + 62) / 63 : 1) * 64;
        }
        const long segs = ((len + sw - 2) / (sw - 1)) * rows;
        return (segs * sw + 63) / 64 * 64;
    };
    switch (kind) {
        case BK_BCC_X: return sides * (ny + 6);
        case BK_BCC_Y: return 2 * (nx + 6) * 3;  // the contiguous axis: one item per ghost cell, both sides
        case BK_BCF_X: return sides * (ny + 7);
        case BK_BCF_Y: return sides * (nx + 7);
        case BK_CELLS1: return (nx + 2) * (ny + 2);
        case BK_PRED_X: return chunked(nx + 3, ny + 4);
        case BK_PRED_Y: return segments(ny + 3, nx + 4);
        case BK_EDGES_P: return (nx + 3) * (ny + 3);
        case BK_FACES: return q.nfc;
        case BK_FLUX_X: return chunked(nx + 1, ny + 2);
        case BK_FLUX_Y: return segments(ny + 1, nx + 2);
        case BK_EDGES: return (nx + 1) * (ny + 1);
        default: return q.nc;  // BK_CELLS
    }
}

static void batch_drop_graph(dg_batch* b) {
    if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->gexec = nullptr; b->graph = nullptr; b->graph_steps = 0;
}

// (Re)build the block table of one kind on the device.
static int batch_table(dg_batch* b, int kind) {
    std::vector<int> patch, first;
    for (int p : b->order) {
        const long items = batch_items(b, b->P[p], kind);
        for (long f = 0; f < items; f += BATCH_THREADS) {
            patch.push_back(p);
            first.push_back((int)f);
        }
    }
    const size_t nb = patch.size();
    if (b->bt_mem[kind] && hipFree(b->bt_mem[kind]) != hipSuccess) return DG_ERR_HIP;
    b->bt_mem[kind] = nullptr;
    if (hipMalloc((void**)&b->bt_mem[kind], sizeof(int) * 2 * std::max<size_t>(nb, 1)) != hipSuccess) return DG_ERR_HIP;
    if (nb) {
        if (hipMemcpy(b->bt_mem[kind], patch.data(), sizeof(int) * nb, hipMemcpyHostToDevice) != hipSuccess) return DG_ERR_HIP;
        if (hipMemcpy(b->bt_mem[kind] + nb, first.data(), sizeof(int) * nb, hipMemcpyHostToDevice) != hipSuccess) return DG_ERR_HIP;
    }
    b->bt[kind] = BlockTable{b->bt_mem[kind], b->bt_mem[kind] + nb, (unsigned)nb};
    return DG_OK;
}

static BatchArgs batch_args(const dg_batch* b, int commit, int cls_filter) {
    BatchArgs a{};
    a.P = b->dP; a.npatch = b->npatch;
    a.W[0] = b->W[0]; a.W[1] = b->W[1]; a.B[0] = b->B[0]; a.B[1] = b->B[1];
    a.Fp = b->Fp; a.F = b->F; a.Bh = b->Bh; a.E0 = b->E0; a.Eh = b->Eh;
    a.parity = b->parity; a.dt_last = b->dt_last; a.steps = b->steps; a.fails = b->fails; a.status = b->status;
    a.err = b->err; a.vmax = b->vmax; a.dt = b->dtp; a.halv = b->halv; a.active = b->active;
    a.vnext = b->vnext; a.cfl_ok = b->cfl_ok;
    a.ph = b->ph;
    a.cfl_coeff = b->d.cfl_coeff; a.tol = b->d.timestep_tol > 0 ? b->d.timestep_tol : 1e-14;
    a.cfl_mode = b->d.cfl_mode; a.bc_kind = b->bc_kind;
    for (int c = 0; c < BATCH_CLASSES; c++) { a.chunk[c] = b->chunk[c]; a.ysw[c] = b->ysw[c]; }
    a.commit = commit; a.cls_filter = cls_filter;
    return a;
}

static void batch_launch(dg_batch* b, int commit, int cls_filter, bool init_only) {
    const BatchArgs a = batch_args(b, commit, cls_filter);
    if (b->d.precision == DG_MIXED) p32::launch_batch_step(a, b->bt, init_only, b->s);
    else p64::launch_batch_step(a, b->bt, init_only, b->s);
}

extern "C" {

int dg_batch_create(const dg_desc* physics, int npatch, const int* nx, const int* ny, const double* dx, const double* dy,
                    int bc_kind, dg_batch** out) {
    if (!physics || !out || npatch < 1 || !nx || !ny || !dx || !dy) return DG_ERR_ARG;
    if (!physics->mhd || !physics->ctu || !physics->muscl || (bc_kind != DG_BC_OUTFLOW && bc_kind != DG_BC_PERIODIC)) return DG_ERR_ARG;
    for (int p = 0; p < npatch; p++) {
        if (nx[p] < 4 || ny[p] < 4 || !(dx[p] > 0) || !(dy[p] > 0)) return DG_ERR_ARG;
        if ((nx[p] + 7L) * (ny[p] + 7L) >= DG_MAX_ELEMS) return DG_ERR_ARG;  // 32-bit indexing within a patch (3 ghosts)
    }
    dg_batch* b = new dg_batch();
    b->d = *physics;
    b->d.dim = 2;
    b->ph = make_phys(b->d);
    b->npatch = npatch;
    b->bc_kind = bc_kind;
    b->P.resize(npatch);
    const int G = 3;
    for (int p = 0; p < npatch; p++) {
        PatchDesc& q = b->P[p];
        q.n[0] = nx[p]; q.n[1] = ny[p];
        q.d[0] = dx[p]; q.d[1] = dy[p];
        q.nc = (long)(nx[p] + 2 * G) * (ny[p] + 2 * G);
        q.nfc = (long)(nx[p] + 1 + 2 * G) * (ny[p] + 1 + 2 * G);
        q.oc = b->cells; q.of = b->faces;
        q.cls = batch_class(nx[p], ny[p]);
        b->cells += q.nc; b->faces += q.nfc;
        b->zones += (long)nx[p] * ny[p];
        b->max_nc = std::max(b->max_nc, q.nc);
        b->max_nfc = std::max(b->max_nfc, q.nfc);
    }
    b->order.resize(npatch);
    for (int p = 0; p < npatch; p++) b->order[p] = p;
    std::stable_sort(b->order.begin(), b->order.end(), [&](int x, int y) { return b->P[x].nc > b->P[y].nc; });
    bool ok = hipStreamCreateWithFlags(&b->s, hipStreamNonBlocking) == hipSuccess;
    auto alloc = [&](void** ptr, size_t bytes) { return ok && hipMalloc(ptr, bytes) == hipSuccess; };
    const size_t cd = sizeof(double) * b->cells, fd = sizeof(double) * b->faces, n = (size_t)npatch;
    ok = alloc((void**)&b->W[0], 8 * cd) && alloc((void**)&b->W[1], 8 * cd) && alloc((void**)&b->B[0], 3 * fd) &&
         alloc((void**)&b->B[1], 3 * fd) && alloc((void**)&b->Fp, 16 * cd) && alloc((void**)&b->F, 16 * cd) &&
         alloc((void**)&b->Bh, 3 * fd) && alloc((void**)&b->E0, 3 * fd) && alloc((void**)&b->Eh, 3 * fd) &&
         alloc((void**)&b->dP, sizeof(PatchDesc) * n) && alloc((void**)&b->parity, sizeof(int) * n) &&
         alloc((void**)&b->dt_last, sizeof(double) * n) && alloc((void**)&b->steps, sizeof(long) * n) &&
         alloc((void**)&b->fails, sizeof(long) * n) && alloc((void**)&b->status, sizeof(int) * n) &&
         alloc((void**)&b->err, sizeof(unsigned long long) * n) && alloc((void**)&b->vmax, sizeof(unsigned long long) * n) &&
         alloc((void**)&b->dtp, sizeof(double) * n) && alloc((void**)&b->halv, sizeof(int) * n) &&
         alloc((void**)&b->active, sizeof(int) * n) && alloc((void**)&b->vnext, sizeof(unsigned long long) * n) &&
         alloc((void**)&b->cfl_ok, sizeof(int) * n) &&
         alloc((void**)&b->stage, sizeof(double) * std::max(8 * b->max_nc, 3 * b->max_nfc));
    ok = ok && hipMemcpyAsync(b->dP, b->P.data(), sizeof(PatchDesc) * n, hipMemcpyHostToDevice, b->s) == hipSuccess;
    // set-up only (never inside a captured graph)
    ok = ok && hipMemsetAsync(b->W[0], 0, 8 * cd, b->s) == hipSuccess && hipMemsetAsync(b->W[1], 0, 8 * cd, b->s) == hipSuccess;
    ok = ok && hipMemsetAsync(b->B[0], 0, 3 * fd, b->s) == hipSuccess && hipMemsetAsync(b->B[1], 0, 3 * fd, b->s) == hipSuccess;
    ok = ok && hipMemsetAsync(b->parity, 0, sizeof(int) * n, b->s) == hipSuccess;
    ok = ok && hipMemsetAsync(b->dt_last, 0, sizeof(double) * n, b->s) == hipSuccess;
    ok = ok && hipMemsetAsync(b->steps, 0, sizeof(long) * n, b->s) == hipSuccess;
    ok = ok && hipMemsetAsync(b->fails, 0, sizeof(long) * n, b->s) == hipSuccess;
    ok = ok && hipMemsetAsync(b->status, 0, sizeof(int) * n, b->s) == hipSuccess;
    ok = ok && hipMemsetAsync(b->halv, 0, sizeof(int) * n, b->s) == hipSuccess;
    ok = ok && hipMemsetAsync(b->active, 0, sizeof(int) * n, b->s) == hipSuccess;
    ok = ok && hipMemsetAsync(b->vnext, 0, sizeof(unsigned long long) * n, b->s) == hipSuccess;
    ok = ok && hipMemsetAsync(b->cfl_ok, 0, sizeof(int) * n, b->s) == hipSuccess;
    ok = ok && hipEventCreate(&b->e0) == hipSuccess && hipEventCreate(&b->e1) == hipSuccess;
    ok = ok && hipStreamSynchronize(b->s) == hipSuccess;
    for (int k = 0; ok && k < BK_COUNT; k++) ok = batch_table(b, k) == DG_OK;
    if (!ok) {
        dg_batch_destroy(b);
        return DG_ERR_HIP;
    }
    *out = b;
    return DG_OK;
}

int dg_batch_destroy(dg_batch* b) {
    if (!b) return DG_OK;
    int rc = DG_OK;
    if (b->s && hipStreamSynchronize(b->s) != hipSuccess) rc = DG_ERR_HIP;
    batch_drop_graph(b);
    void* bufs[] = {b->W[0], b->W[1], b->B[0], b->B[1], b->Fp, b->F, b->Bh, b->E0, b->Eh, b->dP, b->parity, b->dt_last,
                    b->steps, b->fails, b->status, b->err, b->vmax, b->dtp, b->halv, b->active, b->stage, b->vnext, b->cfl_ok};
    for (void* p : bufs)
        if (p && hipFree(p) != hipSuccess) rc = DG_ERR_HIP;
    for (int k = 0; k < BK_COUNT; k++)
        if (b->bt_mem[k] && hipFree(b->bt_mem[k]) != hipSuccess) rc = DG_ERR_HIP;
    if (b->e0 && hipEventDestroy(b->e0) != hipSuccess) rc = DG_ERR_HIP;
    if (b->e1 && hipEventDestroy(b->e1) != hipSuccess) rc = DG_ERR_HIP;
    if (b->s && hipStreamDestroy(b->s) != hipSuccess) rc = DG_ERR_HIP;
    delete b;
    return rc;
}

long dg_batch_zones(const dg_batch* b) { return b ? b->zones : -1; }
int dg_batch_ghosts(const dg_batch* b) { return b ? 3 : -1; }
const char* dg_batch_last_error(const dg_batch* b) { return b ? b->err_msg.c_str() : "null handle"; }
void* dg_batch_stream(dg_batch* b) { return b ? (void*)b->s : nullptr; }

int dg_batch_upload(dg_batch* b, int patch, const double* w, int stride, const double* B) {
    if (!b || patch < 0 || patch >= b->npatch || !w || stride != 8 || !B) return DG_ERR_ARG;
    const PatchDesc& q = b->P[patch];
    BCHK(hipStreamSynchronize(b->s));
    int par = 0;
    BCHK(hipMemcpy(&par, b->parity + patch, sizeof(int), hipMemcpyDeviceToHost));
    BCHK(hipMemcpyAsync(b->stage, w, sizeof(double) * 8 * q.nc, hipMemcpyHostToDevice, b->s));
    p64::launch_aos_to_soa(b->stage, 8, 8, soa(b->W[par] + 8 * q.oc, q.nc), q.nc, b->s);
    BCHK(hipMemcpyAsync(b->stage, B, sizeof(double) * 3 * q.nfc, hipMemcpyHostToDevice, b->s));
    p64::launch_aos_to_soa(b->stage, 3, 3, soa(b->B[par] + 3 * q.of, q.nfc), q.nfc, b->s);
    // a fresh state: clear the patch's counters and any stopped status
    const int zero = 0;
    const long lzero = 0;
    BCHK(hipMemcpyAsync(b->status + patch, &zero, sizeof(int), hipMemcpyHostToDevice, b->s));
    BCHK(hipMemcpyAsync(b->halv + patch, &zero, sizeof(int), hipMemcpyHostToDevice, b->s));
    BCHK(hipMemcpyAsync(b->cfl_ok + patch, &zero, sizeof(int), hipMemcpyHostToDevice, b->s));  // a new state: full CFL next
    BCHK(hipMemcpyAsync(b->steps + patch, &lzero, sizeof(long), hipMemcpyHostToDevice, b->s));
    BCHK(hipMemcpyAsync(b->fails + patch, &lzero, sizeof(long), hipMemcpyHostToDevice, b->s));
    BCHK(hipStreamSynchronize(b->s));
    return DG_OK;
}

int dg_batch_download(dg_batch* b, int patch, double* w, int stride, double* B) {
    if (!b || patch < 0 || patch >= b->npatch || !w || stride != 8) return DG_ERR_ARG;
    const PatchDesc& q = b->P[patch];
    BCHK(hipStreamSynchronize(b->s));
    int par = 0;
    BCHK(hipMemcpy(&par, b->parity + patch, sizeof(int), hipMemcpyDeviceToHost));
    p64::launch_soa_to_aos(soa(b->W[par] + 8 * q.oc, q.nc), 8, b->stage, 8, q.nc, b->s);
    BCHK(hipMemcpyAsync(w, b->stage, sizeof(double) * 8 * q.nc, hipMemcpyDeviceToHost, b->s));
    BCHK(hipStreamSynchronize(b->s));
    if (B) {
        p64::launch_soa_to_aos(soa(b->B[par] + 3 * q.of, q.nfc), 3, b->stage, 3, q.nfc, b->s);
        BCHK(hipMemcpyAsync(B, b->stage, sizeof(double) * 3 * q.nfc, hipMemcpyDeviceToHost, b->s));
        BCHK(hipStreamSynchronize(b->s));
    }
    return DG_OK;
}

int dg_batch_init_B(dg_batch* b) {
    if (!b) return DG_ERR_ARG;
    BCHK(hipMemsetAsync(b->cfl_ok, 0, sizeof(int) * b->npatch, b->s));  // the cells' B change: full CFL next (not captured)
    batch_launch(b, 1, -1, true);
    BCHK(hipGetLastError());
    BCHK(hipStreamSynchronize(b->s));
    return DG_OK;
}

int dg_batch_run(dg_batch* b, int nsteps, int use_graph, double* device_ms) {
    if (!b || nsteps < 0) return DG_ERR_ARG;
    if (nsteps == 0) return DG_OK;
    if (use_graph && (!b->gexec || b->graph_steps != nsteps)) {
        batch_drop_graph(b);
        BCHK(hipStreamBeginCapture(b->s, hipStreamCaptureModeThreadLocal));
        for (int k = 0; k < nsteps; k++) batch_launch(b, 1, -1, false);
        hipGraph_t gr = nullptr;
        BCHK(hipStreamEndCapture(b->s, &gr));
        b->graph = gr;
        BCHK(hipGraphInstantiate(&b->gexec, b->graph, nullptr, nullptr, 0));
        b->graph_steps = nsteps;
    }
    if (device_ms) BCHK(hipEventRecord(b->e0, b->s));
    if (use_graph) {
        BCHK(hipGraphLaunch(b->gexec, b->s));
    } else {
        for (int k = 0; k < nsteps; k++) batch_launch(b, 1, -1, false);
        BCHK(hipGetLastError());
    }
    if (device_ms) {
        BCHK(hipEventRecord(b->e1, b->s));
        BCHK(hipEventSynchronize(b->e1));
        float ms = 0;
        BCHK(hipEventElapsedTime(&ms, b->e0, b->e1));
        *device_ms = ms;
    }
    return DG_OK;
}

int dg_batch_sync(dg_batch* b, dg_status* st) {
    if (!b) return DG_ERR_ARG;
    BCHK(hipStreamSynchronize(b->s));
    std::vector<int> status(b->npatch);
    BCHK(hipMemcpy(status.data(), b->status, sizeof(int) * b->npatch, hipMemcpyDeviceToHost));
    for (int p = 0; p < b->npatch; p++)
        if (status[p]) {
            const char* msg = status[p] == DG_ERR_CFL ? "CFL timestep failed: max signal speed is zero"
                                                       : "Timestep has fallen below minimum. Exiting";
            set_status(st, status[p], p, 0, 0, msg);
            b->err_msg = std::string(msg) + " (patch " + std::to_string(p) + ")";
            return status[p];
        }
    set_status(st, DG_OK, 0, 0, 0, "");
    return DG_OK;
}

int dg_batch_patch_stats(dg_batch* b, int patch, double* dt_last, long* steps, long* fails, int* status) {
    if (!b || patch < 0 || patch >= b->npatch) return DG_ERR_ARG;
    BCHK(hipStreamSynchronize(b->s));
    if (dt_last) BCHK(hipMemcpy(dt_last, b->dt_last + patch, sizeof(double), hipMemcpyDeviceToHost));
    if (steps) BCHK(hipMemcpy(steps, b->steps + patch, sizeof(long), hipMemcpyDeviceToHost));
    if (fails) BCHK(hipMemcpy(fails, b->fails + patch, sizeof(long), hipMemcpyDeviceToHost));
    if (status) BCHK(hipMemcpy(status, b->status + patch, sizeof(int), hipMemcpyDeviceToHost));
    return DG_OK;
}

int dg_batch_set_tiles(dg_batch* b, const int* chunk, int n) {
    if (!b || !chunk || n < 1 || n > BATCH_CLASSES) return DG_ERR_ARG;
    for (int c = 0; c < n; c++)
        if (chunk[c] < 1 || chunk[c] > 4096) return DG_ERR_ARG;
    BCHK(hipStreamSynchronize(b->s));
    for (int c = 0; c < n; c++) b->chunk[c] = chunk[c];
    batch_drop_graph(b);
    int rc = batch_table(b, BK_PRED_X);
    return rc ? rc : batch_table(b, BK_FLUX_X);
}

int dg_batch_get_tiles(const dg_batch* b, int* chunk, int n) {
    if (!b || !chunk || n < 1 || n > BATCH_CLASSES) return DG_ERR_ARG;
    for (int c = 0; c < n; c++) chunk[c] = b->chunk[c];
    return DG_OK;
}

int dg_batch_autotune(dg_batch* b, int reps, int* chunk_out, double* class_ms) {
    if (!b || reps < 1) return DG_ERR_ARG;
    static const int cand[] = {2, 4, 8, 16, 32, 64, 128};
    static const int ycand[] = {64, 32, 16, 0};
    std::vector<int> ncls(BATCH_CLASSES, 0);
    for (const auto& q : b->P) ncls[q.cls]++;
    BCHK(hipStreamSynchronize(b->s));
    batch_drop_graph(b);
    // one class's step (commit off: states and counters untouched), timed
    auto time_class = [&](int c, float* ms) -> int {
        batch_launch(b, 0, c, false);  // warm
        BCHK(hipEventRecord(b->e0, b->s));
        for (int r = 0; r < reps; r++) batch_launch(b, 0, c, false);
        BCHK(hipEventRecord(b->e1, b->s));
        BCHK(hipEventSynchronize(b->e1));
        BCHK(hipEventElapsedTime(ms, b->e0, b->e1));
        return DG_OK;
    };
    for (int c = 0; c < BATCH_CLASSES; c++) {
        if (class_ms) class_ms[c] = 0;
        if (!ncls[c]) continue;
        double best = 1e300;
        int pick = b->chunk[c];
        for (int k : cand) {  // x sweep: faces per walker
            b->chunk[c] = k;
            if (batch_table(b, BK_PRED_X) || batch_table(b, BK_FLUX_X)) return DG_ERR_HIP;
            float ms = 0;
            if (int rc = time_class(c, &ms)) return rc;
            if (ms < best) { best = ms; pick = k; }
        }
        b->chunk[c] = pick;
        if (batch_table(b, BK_PRED_X) || batch_table(b, BK_FLUX_X)) return DG_ERR_HIP;
        int ypick = b->ysw[c];
        best = 1e300;
        for (int w : ycand) {  // y sweep: segment width
            b->ysw[c] = w;
            if (batch_table(b, BK_PRED_Y) || batch_table(b, BK_FLUX_Y)) return DG_ERR_HIP;
            float ms = 0;
            if (int rc = time_class(c, &ms)) return rc;
            if (ms < best) { best = ms; ypick = w; }
        }
        b->ysw[c] = ypick;
        if (class_ms) class_ms[c] = best / reps;
    }
    for (int kind : {BK_PRED_X, BK_FLUX_X, BK_PRED_Y, BK_FLUX_Y})
        if (batch_table(b, kind)) return DG_ERR_HIP;
    if (chunk_out)
        for (int c = 0; c < BATCH_CLASSES; c++) chunk_out[c] = b->chunk[c];
    return DG_OK;
}

int dg_batch_set_ysweep(dg_batch* b, const int* width, int n) {
    if (!b || !width || n < 1 || n > BATCH_CLASSES) return DG_ERR_ARG;
    for (int c = 0; c < n; c++)
        if (width[c] != 64 && width[c] != 32 && width[c] != 16 && width[c] != 0) return DG_ERR_ARG;
    BCHK(hipStreamSynchronize(b->s));
    batch_drop_graph(b);
    for (int c = 0; c < n; c++) b->ysw[c] = width[c];
    if (batch_table(b, BK_PRED_Y) || batch_table(b, BK_FLUX_Y)) return DG_ERR_HIP;
    return DG_OK;
}

int dg_batch_get_ysweep(const dg_batch* b, int* width, int n) {
    if (!b || !width || n < 1 || n > BATCH_CLASSES) return DG_ERR_ARG;
    for (int c = 0; c < n; c++) width[c] = b->ysw[c];
    return DG_OK;
}

}  // extern "C"
