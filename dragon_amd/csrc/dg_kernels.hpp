// dg_kernels.hpp -- launch interface between the host runtime (dg_api.hip)
// and the kernels (dg_kernels.hip).
#pragma once
#include "dg_grid.hpp"
#include "dg_math.hpp"

namespace dg {

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
};

// One boundary pass (cells or faces of one face/axis of one BC element).
struct BCArgs {
    Soa W;
    double* B[3];
    Ext ce, fe;
    int dim, nf, kind, side, axis, corners, gated;  // side: 0 negative, 1 positive, 2 both (periodic)
    double st[8];
    int lo[2], hi[2];                           // transverse ranges
};
void launch_bc(const BCArgs& a, bool faces, hipStream_t s);
void launch_cfl(const StepCtx& c, unsigned long long* out, hipStream_t s);
void launch_step(const StepCtx& c, hipStream_t s);
void ctu_hydro3(const StepCtx& c, hipStream_t s);
void launch_init_B(int dim, Soa W, double* B, Ext ce, Ext fe, hipStream_t s);
void launch_divb(int dim, double* B, Ext ce, Ext fe, const double d[3], unsigned long long* out, hipStream_t s);
void launch_fill_u64(unsigned long long* p, unsigned long long v, int n, hipStream_t s);
void launch_aos_to_soa(const double* a, int stride, int nf, Soa soa, long n, hipStream_t s);
void launch_soa_to_aos(Soa soa, int nf, double* a, int stride, long n, hipStream_t s);

}  // namespace dg
