// dg_kernels.hpp -- launch interface between the host runtime (dg_api.hip)
// and the kernels (dg_kernels.hip).
#pragma once
#include "dg_grid.hpp"
#include "dg_math.hpp"

namespace dg {

// Kernel classes for the optional live per-kernel timing (HIP events recorded
// after each launch on the grid's stream).
enum KernelId {
    K_BC, K_CFL, K_HALF, K_PRED_X, K_PRED_Y, K_PRED_Z, K_EMF_PRED, K_FARADAY_HALF, K_WHALF, K_CTU,
    K_FLUX_X, K_FLUX_Y, K_FLUX_Z, K_EMF, K_FARADAY, K_APPLY, K_COUNT
};
static const char* const kKernelNames[K_COUNT] = {
    "boundary", "cfl", "half_states", "pred_sweep_x", "pred_sweep_y", "pred_sweep_z", "emf_pred", "faraday_half",
    "whalf", "ctu_correct", "flux_sweep_x", "flux_sweep_y", "flux_sweep_z", "emf", "faraday", "apply"};

// Everything one unsplit step needs: geometry, physics switches, the current
// and next state, and the scratch arrays of the stages.
struct StepCtx {
    int dim;
    bool mhd, mixed;
    int ctu, cfl_mode;
    bool fused;           // fused sweeps (no half-state arrays) where available
    int march_chunk;      // faces one thread walks in a strided-axis sweep (0: automatic)
    int march_window;     // 1: FP64 walkers run the register-window form of the FP32 build (a check; same results)
    Ext ce, fe;
    Phys ph;
    double d[3];
    double dt;
    Soa W, Wn;            // current / next cell state
    double* B[3];         // current face B
    double* Bn[3];        // next face B
    double* Bh[3];        // half-step face B (CTU MHD)
    Soa HS[6];            // xL, xR, yL, yR, zL, zR
    Soa F[3];             // corrector (final) fluxes; reference-structured path: both
    Soa Fp[3];            // predictor fluxes of the fused path
    Soa T[5];             // 3D hydro corner fluxes
    double* E0[3];        // corner EMF (predictor, then final)
    double* Eh[3];        // half-step reference EMF
    unsigned long long* err;
    double* cflpart;      // per-block CFL maxima of the new state (apply pass), or null
    void (*mark)(void* user, int id);  // per-kernel timing hook (null when off)
    void* mark_user;
};
inline void mark(const StepCtx& c, int id) {
    if (c.mark) c.mark(c.mark_user, id);
}

// One boundary pass (cells or faces of one face/axis of one BC element).
struct BCArgs {
    Soa W;
    double* B[3];
    Ext ce, fe;
    int dim, nf, kind, side, axis, corners, gated;  // side: 0 negative, 1 positive, 2 both (periodic)
    double st[8];
    int lo[2], hi[2];                           // transverse ranges
};
// Batched 2D MHD patches (CTU + CT).  Every stage of the step is one launch
// over all patches: a block table per work kind maps each workgroup to a
// patch and the first work item it covers.  Pools hold every patch back to
// back: patch p's cell field f of pool X is at X + nfield*oc + f*nc; face
// component c at X + 3*of + c*nfc.
struct PatchDesc {
    int n[2];           // interior cells
    double d[2];        // dx, dy
    long oc, nc;        // cell offset / count (ghosts included)
    long of, nfc;       // face offset / count
    int cls;            // size class (tile table index)
};
enum { BATCH_CLASSES = 4, BATCH_THREADS = 256 };
// per-patch status codes (the values of DG_ERR_CFL / DG_ERR_DT_UNDERFLOW)
enum { BATCH_ERR_CFL = 5, BATCH_ERR_DT = 6 };
// work kinds of a batched step (items per patch in dg_api.hip: batch_items)
enum BatchKind {
    BK_BCC_X, BK_BCC_Y, BK_BCF_X, BK_BCF_Y,  // boundary passes: cells / faces, x then y, both sides
    BK_CELLS1,                               // cells [-1, n+1)^2: CFL, half-step state
    BK_PRED_X, BK_PRED_Y,                    // predictor sweeps (x: faces-per-thread chunks; y: 63 faces per wavefront)
    BK_EDGES_P,                              // edges [-1, n+2)^2: predictor EMF
    BK_FACES,                                // every face: Faraday
    BK_FLUX_X, BK_FLUX_Y,                    // corrector sweeps
    BK_EDGES,                                // edges [0, n+1)^2: EMF
    BK_CELLS,                                // every cell: apply, initialize_B_fields
    BK_COUNT
};
struct BlockTable {
    const int* patch;   // per workgroup: patch
    const int* first;   // per workgroup: first work item (a multiple of BATCH_THREADS)
    unsigned nblocks;
};
struct BatchArgs {
    const PatchDesc* P;
    int npatch;
    double* W[2];              // cell state pools (8 fields), double-buffered
    double* B[2];              // face-B pools, double-buffered
    double* Fp;                // predictor fluxes, 2 axes x 8 fields
    double* F;                 // corrector fluxes, 2 axes x 8 fields
    double* Bh;                // half-step faces
    double* E0;                // corner EMF
    double* Eh;                // half-step reference EMF
    int* parity;               // per patch: index of the current buffer
    double* dt_last;           // per patch: dt of the last accepted step
    long* steps;               // per patch: accepted steps
    long* fails;               // per patch: rejected attempts
    int* status;               // per patch: 0, or the DG_* code that stopped it
    unsigned long long* err;   // per patch failure key of the current attempt
    unsigned long long* vmax;  // per patch CFL signal speed (FP64 bits)
    unsigned long long* vnext; // per patch: the apply pass's maximum over the interior of the state it wrote (FP64 bits)
    int* cfl_ok;               // per patch: vnext describes the current state (its last attempt was committed)
    double* dt;                // per patch dt of the current attempt
    int* halv;                 // per patch halvings since the last accepted step
    int* active;               // per patch: steps in this launch sequence
    Phys ph;
    double cfl_coeff, tol;
    int cfl_mode, bc_kind;     // bc_kind: 0 outflow, 1 periodic (all faces, corner ghosts)
    int chunk[BATCH_CLASSES];  // x-sweep faces per thread, per size class
    int ysw[BATCH_CLASSES];    // y-sweep segment width in lanes (64, 32, 16: 63, 31, 15 faces), per size class
    int commit;                // 0: run the step, keep every patch's state and counters (tuning)
    int cls_filter;            // -1: all patches, else only this size class
};

// The kernels are compiled twice (dg_kernels.hip): p64 is the FP64 drop-in
// (no FMA contraction, correctly rounded division and square root); p32 holds
// the mixed form's kernels (FP32 reconstruction and Riemann solves with
// contraction and fast FP32 division / square root; FP64 state, accumulation,
// pressure recovery and CT).  Separate namespaces keep every kernel symbol of
// one build out of the other.
#define DG_LAUNCH_INTERFACE                                                                                     \
    void launch_batch_step(const BatchArgs& a, const BlockTable* bt, bool init_only, hipStream_t s);           \
    void launch_step(const StepCtx& c, hipStream_t s);                                                          \
    void launch_split_axis(const StepCtx& c, int ax, Soa in, Soa out, hipStream_t s);  /* one split sweep (hydro) */ \
    void launch_bc(const BCArgs& a, bool faces, hipStream_t s);                                                 \
    void launch_cfl(const StepCtx& c, unsigned long long* out, hipStream_t s);                                  \
    void launch_cfl_tail(const StepCtx& c, long nparts, unsigned long long* out, hipStream_t s);                \
    void launch_init_B(int dim, Soa W, double* B, Ext ce, Ext fe, hipStream_t s);                               \
    void launch_divb(int dim, double* B, Ext ce, Ext fe, const double d[3], unsigned long long* out, hipStream_t s); \
    void launch_fill_u64(unsigned long long* p, unsigned long long v, int n, hipStream_t s);                    \
    void launch_aos_to_soa(const double* a, int stride, int nf, Soa soa, long n, hipStream_t s);                \
    void launch_soa_to_aos(Soa soa, int nf, double* a, int stride, long n, hipStream_t s);
namespace p64 { DG_LAUNCH_INTERFACE }
namespace p32 { DG_LAUNCH_INTERFACE }

}  // namespace dg
