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
// Batched 2D MHD patches (CTU + CT), one workgroup per patch per step.
// Pools hold every patch back to back: patch p's cell field f of pool X is
// at X + nfield*oc + f*nc; face component c at X + 3*of + c*nfc.
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
struct BatchArgs {
    const PatchDesc* P;
    const int* order;          // processing order: largest patches first
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
    long* fails;               // per patch: rejected attempts (dt halvings)
    int* status;               // per patch: 0, or the DG_* code of the last step that could not complete
    unsigned long long* err;   // per patch failure key of the current attempt
    unsigned* queue;           // [0] next index into order, [1] finished workgroups
    Phys ph;
    double cfl_coeff, tol;
    int cfl_mode, bc_kind;     // bc_kind: 0 outflow, 1 periodic (all faces, corner ghosts)
    int chunk[BATCH_CLASSES];  // x-sweep faces per thread, per size class
    int commit;                // 0: run the step but keep the old state (tuning)
    int init_only;             // 1: boundary + cell B from faces (initialize_B_fields) only
};
// The kernels are compiled twice (dg_kernels.hip): p64 is the FP64 drop-in
// (no FMA contraction, correctly rounded division and square root); p32 holds
// the mixed form's kernels (FP32 reconstruction and Riemann solves with
// contraction and fast FP32 division / square root; FP64 state, accumulation,
// pressure recovery and CT).  Separate namespaces keep every kernel symbol of
// one build out of the other.
#define DG_LAUNCH_INTERFACE                                                                                     \
    void launch_batch(const BatchArgs& a, int blocks, hipStream_t s);                                            \
    int batch_blocks_per_cu();                                                                                  \
    void launch_step(const StepCtx& c, hipStream_t s);                                                          \
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
