// dg_kernels.hip -- CDNA4 kernels for DRAGON's unsplit Godunov step.
//
// One kernel per loop nest of the reference step, each over exactly the index
// range the reference uses, reading and writing SoA FP64 fields in HBM.  The
// ranges are what make the FP64 form bit-compatible with the CPU path; the
// stage order (enforced by stream order) is the reference's.
//
// Kernels are templated on the dimension D, on MHD, and on the arithmetic type
// R of reconstruction and Riemann solves (double: drop-in; float: the mixed
// form, whose state, accumulation, pressure recovery and CT stay double).
#include <algorithm>
#include <type_traits>

#include "dg_kernels.hpp"

#ifndef DG_PREC
#define DG_PREC 64
#endif
#if DG_PREC == 32
#define DG_NS p32
#else
#define DG_NS p64
#endif

namespace dg {
namespace DG_NS {
#if DG_PREC == 32
using RT = float;   // arithmetic type of reconstruction and Riemann solves
using FT = float;   // storage type of the flux arrays (FP32 Riemann fluxes lose nothing)
using ET = float;   // storage type of the corner and reference EMF arrays (the arithmetic on them stays FP64)
#else
using RT = double;
using FT = double;
using ET = double;
#endif
using FSoa = SoaT<FT>;
__host__ __device__ inline FSoa fs(const Soa& a) { return FSoa{reinterpret_cast<FT*>(a.p), a.stride}; }
__host__ __device__ inline ET* et(double* p) { return reinterpret_cast<ET*>(p); }
// A stored element widened to FP64 before any arithmetic
// Element i of an array of the grid, addressed as the (uniform) base plus a
// 32-bit byte offset: every array holds fewer than DG_MAX_ELEMS elements.
template <class T> __device__ __forceinline__ const T& ld32(const T* p, long i) {
    return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(p) + (unsigned)((unsigned)i * (unsigned)sizeof(T)));
}
template <class T> __device__ __forceinline__ double ed(const T* p, long i) { return double(ld32(p, i)); }
void ctu_hydro3(const StepCtx& c, hipStream_t s);

#define KB __launch_bounds__(256)
// Occupancy target of the flux sweeps (waves per SIMD; 0 = compiler's choice).
#ifndef DG_SWEEP_WAVES
#define DG_SWEEP_WAVES 0
#endif
#if DG_SWEEP_WAVES > 0
#define KS __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(DG_SWEEP_WAVES, DG_SWEEP_WAVES)))
#else
#define KS KB
#endif

__device__ __forceinline__ long tid() { return (long)blockIdx.x * blockDim.x + threadIdx.x; }
// Workgroups are dealt to the 8 XCDs in turn.  For passes over consecutive
// points that read the next row's fluxes or edges, tid_xcd gives the workgroups
// of each group of 32 (in launch order) to XCDs so that one XCD holds 4
// consecutive blocks of cells (about two rows of a 512-cell line): the
// neighbour rows a block reads are then mostly in its own XCD's L2.  A
// permutation of the block indices within whole groups of 32; the last,
// partial group keeps its order.
__device__ __forceinline__ unsigned block_xcd() {
    constexpr unsigned R = 4, G = 8 * R;
    unsigned b = blockIdx.x;
    if (b < (gridDim.x & ~(G - 1))) b = (b & ~(G - 1)) | ((b & 7u) * R) | ((b >> 3) & (R - 1));
    return b;
}
__device__ __forceinline__ long tid_xcd() { return (long)block_xcd() * blockDim.x + threadIdx.x; }

template <class R> __device__ __forceinline__ P<R> load_p(const Soa& a, long m, bool mhd) {
    P<R> w;
    w.rho = R(a(F_RHO, m)); w.p = R(a(F_P, m));
    w.vx = R(a(F_VX, m)); w.vy = R(a(F_VY, m)); w.vz = R(a(F_VZ, m));
    if (mhd) { w.Bx = R(a(F_BX, m)); w.By = R(a(F_BY, m)); w.Bz = R(a(F_BZ, m)); }
    else { w.Bx = w.By = w.Bz = R(0); }
    return w;
}
// Streaming (non-temporal) store of the committed state: the apply pass
// writes each cell once and nothing reads it before the next step, so the
// stores need not hold cache lines (apply 9.1 -> 8.8 ms FP64, 6.9 -> 6.5 ms
// mixed at 512^3).
__device__ __forceinline__ void st_nt(const Soa& a, int f, long m, double v) { __builtin_nontemporal_store(v, &a(f, m)); }
__device__ __forceinline__ void store_p_nt(const Soa& a, long m, const P<double>& w, bool mhd) {
    st_nt(a, F_RHO, m, w.rho); st_nt(a, F_P, m, w.p);
    st_nt(a, F_VX, m, w.vx); st_nt(a, F_VY, m, w.vy); st_nt(a, F_VZ, m, w.vz);
    if (mhd) { st_nt(a, F_BX, m, w.Bx); st_nt(a, F_BY, m, w.By); st_nt(a, F_BZ, m, w.Bz); }
}
template <class R> __device__ __forceinline__ void store_p(const Soa& a, long m, const P<R>& w, bool mhd) {
    a(F_RHO, m) = double(w.rho); a(F_P, m) = double(w.p);
    a(F_VX, m) = double(w.vx); a(F_VY, m) = double(w.vy); a(F_VZ, m) = double(w.vz);
    if (mhd) { a(F_BX, m) = double(w.Bx); a(F_BY, m) = double(w.By); a(F_BZ, m) = double(w.Bz); }
}
// Flux records use the U field order: rho, E, mx, my, mz, Bx, By, Bz.
template <class T> __device__ __forceinline__ U<double> load_u(const SoaT<T>& a, long m, bool mhd) {
    U<double> u;
    u.rho = a(0, m); u.E = a(1, m); u.mx = a(2, m); u.my = a(3, m); u.mz = a(4, m);
    if (mhd) { u.Bx = a(5, m); u.By = a(6, m); u.Bz = a(7, m); } else { u.Bx = u.By = u.Bz = 0; }
    return u;
}
// An MHD flux record along axis AX with its normal-B component taken as 0
// instead of loaded: every MHD solver returns it as exactly +0 (HLL, HLLE and
// HLLD average the normal field of both states before the solve, and the
// physical flux sets it to zero), so the sums it enters are unchanged.
template <int AX, class T> __device__ __forceinline__ U<double> load_u_nb(const SoaT<T>& a, long m) {
    U<double> u;
    u.rho = a(0, m); u.E = a(1, m); u.mx = a(2, m); u.my = a(3, m); u.mz = a(4, m);
    u.Bx = AX == 0 ? 0.0 : double(a(5, m));
    u.By = AX == 1 ? 0.0 : double(a(6, m));
    u.Bz = AX == 2 ? 0.0 : double(a(7, m));
    return u;
}
template <class R, class T> __device__ __forceinline__ void store_u(const SoaT<T>& a, long m, const U<R>& u, bool mhd) {
    a(0, m) = T(u.rho); a(1, m) = T(u.E); a(2, m) = T(u.mx); a(3, m) = T(u.my); a(4, m) = T(u.mz);
    if (mhd) { a(5, m) = T(u.Bx); a(6, m) = T(u.By); a(7, m) = T(u.Bz); }
}
__device__ __forceinline__ void cross(double ux, double uy, double uz, double vx, double vy, double vz, double& wx, double& wy, double& wz) {
    wx = uy * vz - uz * vy;  // Arithmetic.cpp:359-365
    wy = uz * vx - ux * vz;
    wz = ux * vy - uy * vx;
}

// ===================================================================== BC
// One thread per transverse index of a face; it runs the reference's ghost
// loop (g = 1..ng) in order, since face-B extrapolation reads the ghost it
// wrote one iteration earlier.  Cell passes and face passes are separate
// launches, issued in the reference's order.


__device__ __forceinline__ long cidx(const Ext& e, int axis, int a, int t0, int t1) {
    // a along `axis`, (t0, t1) the other two dims in increasing order
    int i, j, k;
    if (axis == 0) { i = a; j = t0; k = t1; }
    else if (axis == 1) { i = t0; j = a; k = t1; }
    else { i = t0; j = t1; k = a; }
    return e.at(i, j, k);
}

__device__ void copy_cell(const Soa& W, long dst, long src, int nf) {
    for (int f = 0; f < nf; f++) W(f, dst) = W(f, src);
}

// Ghost cell g (1..ng) on side `side` of the line (t0, t1).
__device__ __forceinline__ void bc_cell(const BCArgs& a, int t0, int t1, int g, int side) {
    const Ext& e = a.ce;
    const int n = e.n[a.axis];
    const int nf = a.nf;
    const int vf = F_VX + a.axis, bfn = F_BX + a.axis;
    {
        if (a.kind == 1) {  // Periodic (Periodic.cpp)
            if (side == 0) copy_cell(a.W, cidx(e, a.axis, -g, t0, t1), cidx(e, a.axis, n - g, t0, t1), nf);
            else copy_cell(a.W, cidx(e, a.axis, n - 1 + g, t0, t1), cidx(e, a.axis, g - 1, t0, t1), nf);
            return;
        }
        const long dst = cidx(e, a.axis, side == 0 ? -g : n - 1 + g, t0, t1);
        if (a.kind == 0) {  // Outflow (+gated) (Outflow.cpp)
            copy_cell(a.W, dst, cidx(e, a.axis, side == 0 ? 0 : n - 1, t0, t1), nf);
            if (a.gated) {
                double v = a.W(vf, dst);
                if (side == 0 ? v > 0 : v < 0) a.W(vf, dst) = 0;
            }
        } else if (a.kind == 2) {  // Reflective (Reflective.cpp)
            copy_cell(a.W, dst, cidx(e, a.axis, side == 0 ? g - 1 : n - g, t0, t1), nf);
            a.W(vf, dst) *= -1;
            if (nf == 8) {
                for (int f = F_BX; f <= F_BZ; f++) a.W(f, dst) *= -1;
                a.W(bfn, dst) *= -1;
            }
        } else if (a.kind == 3) {  // Fixed (Fixed.cpp)
            for (int f = 0; f < nf; f++) a.W(f, dst) = a.st[f];
        }
    }
}
// One thread per line, the reference's g loop in order (periodic: both
// sides per g).
__device__ __forceinline__ void bc_cells_body(const BCArgs& a, long t) {
    const int n1 = a.hi[0] - a.lo[0], n2 = a.hi[1] - a.lo[1];
    if (t >= (long)n1 * n2) return;
    const int t0 = a.lo[0] + (int)(t / n2), t1 = a.lo[1] + (int)(t % n2);
    for (int g = 1; g <= a.ce.g; g++) {
        if (a.kind == 1) { bc_cell(a, t0, t1, g, 0); bc_cell(a, t0, t1, g, 1); }
        else bc_cell(a, t0, t1, g, a.side);
    }
}
KB __global__ void k_bc_cells(BCArgs a) { bc_cells_body(a, tid()); }
// The same with one thread per ghost cell, for the contiguous axis: the
// cells a pass writes are distinct and, when the axis holds at least ng
// cells, every cell it reads is interior, so the g loop has no order to
// keep.  Neighbouring lanes then write neighbouring ghost cells of a line
// instead of one cell per line each.
KB __global__ void k_bc_cells_g(BCArgs a) {
    const int ng = a.ce.g, per = a.kind == 1 ? 2 * ng : ng;
    const int n1 = a.hi[0] - a.lo[0], n2 = a.hi[1] - a.lo[1];
    long t = tid();
    if (t >= (long)n1 * n2 * per) return;
    const int slot = (int)(t % per);
    t /= per;
    const int t0 = a.lo[0] + (int)(t / n2), t1 = a.lo[1] + (int)(t % n2);
    bc_cell(a, t0, t1, slot % ng + 1, a.kind == 1 ? slot / ng : a.side);
}

__device__ __forceinline__ void fget(double* const B[3], long m, double v[3]) { v[0] = B[0][m]; v[1] = B[1][m]; v[2] = B[2][m]; }
__device__ __forceinline__ void fset(double* const B[3], long m, const double v[3]) { B[0][m] = v[0]; B[1][m] = v[1]; B[2][m] = v[2]; }

__device__ __forceinline__ void bc_faces_body(const BCArgs& a, long t) {
    const int n1 = a.hi[0] - a.lo[0], n2 = a.hi[1] - a.lo[1];
    if (t >= (long)n1 * n2) return;
    const int t0 = a.lo[0] + (int)(t / n2), t1 = a.lo[1] + (int)(t % n2);
    const Ext& e = a.fe;
    const int n = a.ce.n[a.axis], ng = e.g, ax = a.axis;
    auto at = [&](int q) { return cidx(e, ax, q, t0, t1); };
    double v[3], u[3];
    if (a.kind == 1) {  // Periodic
        for (int g = 1; g <= ng; g++) {
            fget(a.B, at(n - g), v); fset(a.B, at(-g), v);
            fget(a.B, at(g - 1), v); fset(a.B, at(n - 1 + g), v);
        }
        return;
    }
    if (a.kind == 0) {  // Outflow: copy transverse, extrapolate normal
        if (a.side == 0) {
            for (int g = 1; g <= ng; g++) {
                fget(a.B, at(0), v);
                v[ax] = 2 * a.B[ax][at(-g + 1)] - a.B[ax][at(-g + 2)];
                fset(a.B, at(-g), v);
            }
        } else {
            for (int c = 0; c < 3; c++) if (c != ax) a.B[c][at(n)] = a.B[c][at(n - 1)];
            for (int g = 1; g <= ng; g++) {
                fget(a.B, at(n), v);
                v[ax] = 2 * a.B[ax][at(n + g - 1)] - a.B[ax][at(n + g - 2)];
                fset(a.B, at(n + g), v);
            }
        }
    } else if (a.kind == 2) {  // Reflective: invert, copy normal from the mirror face
        if (a.side == 0) {
            for (int g = 1; g <= ng; g++) {
                fget(a.B, at(g - 1), u);
                for (int c = 0; c < 3; c++) v[c] = u[c] * -1;
                v[ax] = a.B[ax][at(g)];
                fset(a.B, at(-g), v);
            }
        } else {
            fget(a.B, at(n - 1), u);
            for (int c = 0; c < 3; c++) v[c] = u[c] * -1;
            fset(a.B, at(n), v);
            for (int g = 1; g <= ng; g++) {
                fget(a.B, at(n - g - 1), u);
                for (int c = 0; c < 3; c++) v[c] = u[c] * -1;
                v[ax] = a.B[ax][at(n - g)];
                fset(a.B, at(n + g), v);
            }
        }
    } else if (a.kind == 3) {  // Fixed
        const double s[3] = {a.st[F_BX], a.st[F_BY], a.st[F_BZ]};
        if (a.side == 0) {
            a.B[ax][at(0)] = s[ax];
            for (int g = 1; g <= ng; g++) fset(a.B, at(-g), s);
        } else {
            for (int g = 0; g < ng; g++) fset(a.B, at(n + g), s);
        }
    }
}
KB __global__ void k_bc_faces(BCArgs a) { bc_faces_body(a, tid()); }
// Periodic face pass on the contiguous axis, one thread per ghost face:
// with at least ng cells along the axis every face it reads is interior.
KB __global__ void k_bc_faces_periodic_g(BCArgs a) {
    const int ng = a.fe.g, per = 2 * ng;
    const int n1 = a.hi[0] - a.lo[0], n2 = a.hi[1] - a.lo[1];
    long t = tid();
    if (t >= (long)n1 * n2 * per) return;
    const int slot = (int)(t % per);
    t /= per;
    const int t0 = a.lo[0] + (int)(t / n2), t1 = a.lo[1] + (int)(t % n2);
    const int g = slot % ng + 1, n = a.ce.n[a.axis];
    auto at = [&](int q) { return cidx(a.fe, a.axis, q, t0, t1); };
    double v[3];
    if (slot < ng) { fget(a.B, at(n - g), v); fset(a.B, at(-g), v); }
    else { fget(a.B, at(g - 1), v); fset(a.B, at(n - 1 + g), v); }
}

// ==================================================================== CFL
// Max signal speed over cells [-1, n+1)^D (CFL.cpp:113-162); wave64
// shuffles then one atomicMax per block on the (non-negative) FP64 bits.
template <int D, bool MHD>
__device__ __forceinline__ double cfl_speed_w(const P<double>& w, const Phys& ph, int mode, double dx, double dy, double dz) {
    {
        double a = MHD ? c_fast_max(w, ph) : cs(w, ph);
        double s = 0;
        if (D == 1) {
            s = dabs(w.vx) + a;
        } else if (mode == 0) {  // CFL_MAX
            if (dx > 1e-14) s = (dabs(w.vx) + a) / dx;
            if (dy > 1e-14) s = dmax(s, (dabs(w.vy) + a) / dy);
            if (D == 3 && dz > 1e-14) s = dmax(s, (dabs(w.vz) + a) / dz);
        } else if (mode == 1) {  // CFL_ADD
            if (dx > 1e-14) s = (dabs(w.vx) + a) / dx;
            if (dy > 1e-14) s += (dabs(w.vy) + a) / dy;
            if (D == 3 && dz > 1e-14) s += (dabs(w.vz) + a) / dz;
        } else {  // power mean
            const double p = mode;
            if (dx > 1e-14) s = pow((dabs(w.vx) + a) / dx, p);
            if (dy > 1e-14) s += pow((dabs(w.vy) + a) / dy, p);
            if (D == 3 && dz > 1e-14) s += pow((dabs(w.vz) + a) / dz, p);
            s = pow(s, 1.0 / p);
        }
        return s;
    }
}
template <int D, bool MHD>
__device__ __forceinline__ double cfl_speed_at(const Soa& W, const Ext& ce, int i, int j, int k, const Phys& ph, int mode, double dx,
                                               double dy, double dz) {
    return cfl_speed_w<D, MHD>(load_p<double>(W, ce.at(i, j, k), MHD), ph, mode, dx, dy, dz);
}
// Block max of per-thread values into part[blockIdx.x] (plain store, no atomics).
__device__ __forceinline__ void block_max_store(double m, double* part) {
    for (int off = 32; off > 0; off >>= 1) m = dmax(m, __shfl_xor(m, off));
    __shared__ double wmax[4];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) wmax[wave] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        double b = wmax[0];
        for (int q = 1; q < (int)(blockDim.x >> 6); q++) b = dmax(b, wmax[q]);
        part[blockIdx.x] = b;
    }
}
// CFL of the committed state from the maxima the apply pass left over the
// interior plus a scan of the ghost layer 1 (the shell of [-1, n+1)^D).
template <int D, bool MHD>
KB __global__ void k_cfl_tail(Soa W, Ext ce, const double* part, long nparts, Phys ph, int mode, double dx, double dy, double dz,
                              unsigned long long* out) {
    const long n0 = ce.n[0], n1 = ce.n[1], n2 = ce.n[2];
    // shell slabs: x faces (i = -1, n0) over the full transverse range, then y, then z faces inside
    const long sx = D == 3 ? (n1 + 2) * (n2 + 2) : (D == 2 ? n1 + 2 : 1);
    const long sy = D == 3 ? n0 * (n2 + 2) : (D == 2 ? n0 : 0);
    const long sz = D == 3 ? n0 * n1 : 0;
    const long shell = 2 * (sx + sy + sz);
    const long stride = (long)gridDim.x * blockDim.x;
    double m = 0;
    for (long t = tid(); t < nparts + shell; t += stride) {
        if (t < nparts) { m = dmax(m, part[t]); continue; }
        long r = t - nparts;
        int i = 0, j = 0, k = 0;
        if (r < 2 * sx) {
            i = r < sx ? -1 : (int)n0;
            r %= sx;
            if (D == 3) { j = (int)(r / (n2 + 2)) - 1; k = (int)(r % (n2 + 2)) - 1; }
            else if (D == 2) j = (int)r - 1;
        } else if (r < 2 * (sx + sy)) {
            r -= 2 * sx;
            j = r < sy ? -1 : (int)n1;
            r %= sy;
            if (D == 3) { i = (int)(r / (n2 + 2)); k = (int)(r % (n2 + 2)) - 1; }
            else i = (int)r;
        } else {
            r -= 2 * (sx + sy);
            k = r < sz ? -1 : (int)n2;
            r %= (sz > 0 ? sz : 1);  // 3D only: the shell has no z faces below 3D
            i = (int)(r / n1); j = (int)(r % n1);
        }
        m = dmax(m, cfl_speed_at<D, MHD>(W, ce, i, j, k, ph, mode, dx, dy, dz));
    }
    for (int off = 32; off > 0; off >>= 1) m = dmax(m, __shfl_xor(m, off));
    __shared__ double wmax[4];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) wmax[wave] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        double b = wmax[0];
        for (int q = 1; q < (int)(blockDim.x >> 6); q++) b = dmax(b, wmax[q]);
        if (b > 0) atomicMax(out, (unsigned long long)__double_as_longlong(b));
    }
}
template <int D, bool MHD>
KB __global__ void k_cfl(Soa W, Ext ce, Box box, Phys ph, int mode, double dx, double dy, double dz, unsigned long long* out) {
    double m = 0;
    int i, j, k;
    const long cnt = box.count(), stride = (long)gridDim.x * blockDim.x;
    for (long t = tid(); t < cnt; t += stride)  // grid-stride: few blocks, few atomics
        if (box.point(t, i, j, k)) m = dmax(m, cfl_speed_at<D, MHD>(W, ce, i, j, k, ph, mode, dx, dy, dz));  // NaN never wins
    for (int off = 32; off > 0; off >>= 1) m = dmax(m, __shfl_xor(m, off));
    __shared__ double part[4];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) part[wave] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        double b = part[0];
        for (int q = 1; q < (int)(blockDim.x >> 6); q++) b = dmax(b, part[q]);
        if (b > 0) atomicMax(out, (unsigned long long)__double_as_longlong(b));
    }
}

// ============================================================ half states
// computeHalfStates_X/Y/Z + copyFaceFields (Unsplit.cpp:292-436,
// B_fields.cpp:166-216): every cell, every axis.  Cells at the outer ghost
// layer along an axis get the first-order copy.
template <int D, bool MHD, class R>
KB __global__ void k_half(Soa W, double* Bx, double* By, double* Bz, Ext ce, Ext fe, Soa HS0, Soa HS1, Soa HS2, Soa HS3,
                          Soa HS4, Soa HS5, double dt_dx, double dt_dy, double dt_dz, Phys ph) {
    Box box = make_box(D, -ce.g, ce.n[0] + ce.g, -ce.g, ce.n[1] + ce.g, -ce.g, ce.n[2] + ce.g);
    int i, j, k;
    if (!box.point(tid(), i, j, k)) return;
    const long m = ce.at(i, j, k);
    const int g = ce.g;
    double bx = 0, by = 0, bz = 0;
    if (MHD && D >= 2) {
        bx = (Bx[fe.at(i + 1, j, k)] - Bx[fe.at(i, j, k)]) * dt_dx;
        by = (By[fe.at(i, j + 1, k)] - By[fe.at(i, j, k)]) * dt_dy;
        if (D == 3) bz = (Bz[fe.at(i, j, k + 1)] - Bz[fe.at(i, j, k)]) * dt_dz;
    }
    const int idx[3] = {i, j, k};
    const P<R> wc = load_p<R>(W, m, MHD);
#pragma unroll
    for (int ax = 0; ax < D; ax++) {
        P<R> L, Rr;
        const bool inner = idx[ax] >= -g + 1 && idx[ax] < ce.n[ax] + g - 1;
        if (inner) {
            const long off = ce.s[ax];
            P<R> wl = load_p<R>(W, m - off, MHD), wr = load_p<R>(W, m + off, MHD);
            if (ax == 0) {
                muscl<MHD, R>(wl, L, wc, Rr, wr, R(dt_dx), R(bx), R(by), R(bz), ph);
            } else if (ax == 1) {
                P<R> l, r;
                muscl<MHD, R>(swapXY(wl), l, swapXY(wc), r, swapXY(wr), R(dt_dy), R(by), R(bx), R(bz), ph);
                L = swapXY(l); Rr = swapXY(r);
            } else {
                P<R> l, r;
                muscl<MHD, R>(swapXZ(wl), l, swapXZ(wc), r, swapXZ(wr), R(dt_dz), R(bz), R(by), R(bx), ph);
                L = swapXZ(l); Rr = swapXZ(r);
            }
        } else {
            L = wc; Rr = wc;
        }
        if (MHD && D >= 2) {  // normal B of the half states := face B
            if (ax == 0) { L.Bx = R(Bx[fe.at(i, j, k)]); Rr.Bx = R(Bx[fe.at(i + 1, j, k)]); }
            else if (ax == 1) { L.By = R(By[fe.at(i, j, k)]); Rr.By = R(By[fe.at(i, j + 1, k)]); }
            else { L.Bz = R(Bz[fe.at(i, j, k)]); Rr.Bz = R(Bz[fe.at(i, j, k + 1)]); }
        }
        // the FP32 form stores half states in FP64 arrays; exact for floats
        if (ax == 0) { store_p<R>(HS0, m, L, MHD); store_p<R>(HS1, m, Rr, MHD); }
        else if (ax == 1) { store_p<R>(HS2, m, L, MHD); store_p<R>(HS3, m, Rr, MHD); }
        else { store_p<R>(HS4, m, L, MHD); store_p<R>(HS5, m, Rr, MHD); }
    }
}

// ================================================================= fluxes
// computeFlux_X/Y/Z (Unsplit.cpp:210-254): F[c] = Riemann(R[c - e], L[c]).
template <int D, bool MHD, class R, int AX>
KB __global__ void k_flux(Soa HL, Soa HR, FSoa F, Ext ce, Box box, double dt_d, Phys ph, unsigned long long* err, int stage) {
    int i, j, k;
    if (!box.point(tid(), i, j, k)) return;
    const long m = ce.at(i, j, k);
    P<R> l = load_p<R>(HR, m - ce.s[AX], MHD), r = load_p<R>(HL, m, MHD);
    if (AX == 1) { l = swapXY(l); r = swapXY(r); }
    if (AX == 2) { l = swapXZ(l); r = swapXZ(r); }
    int fail = RF_OK;
    U<R> f = riemann_x<MHD, R>(l, r, R(dt_d), ph, &fail);
    if (fail) report(err, stage, (unsigned long long)m, fail == RF_BAD_INPUT ? E_RIEMANN_INPUT : E_RIEMANN_SOLUTION);
    if (AX == 1) f = swapXY(f);
    if (AX == 2) f = swapXZ(f);
    store_u<R>(F, m, f, MHD);
}

// ========================================================= fused sweeps
// Flux along one axis with the half states rebuilt in registers: MUSCL
// (+ normal-B override) and, for the corrector pass, the CTU correction of
// the cell (CTU.cpp:65-129 / 162-202 for MHD, 264-278 for hydro 2D).  The
// half-state arrays of the reference never touch HBM.  Strided axes walk a
// chunk of faces per thread, carrying the right state of the previous cell in
// registers; the contiguous axis hands it to the next lane with a wave64
// shuffle (63 faces per wavefront).
struct SweepArgs {
    Soa W;
    const double* B[3];
    const ET* Bh[3];       // half-step face B (ET storage)
    const ET* E0[3];
    FSoa Fp[3];            // predictor fluxes (corrector pass input)
    FSoa Fo;               // output fluxes of this axis
    Ext ce, fe;
    int lo[3], hi[3];      // faces [lo, hi) per dim; face f of axis AX sits between cells f-1 and f
    double dt_d[3];
    Phys ph;
    unsigned long long* err;
    int stage, chunk;
    ET* eref[3];           // MHD predictor sweep along the contiguous axis: the cells' reference EMF cross(B, v) (else unused)
};
// The MHD predictor sweep along the contiguous axis also writes each rebuilt
// cell's reference EMF, cross(B, v) of the state it holds: its cells cover
// the predictor EMF's ([-2, n+1] in every dimension). In the FP32 form the
// EMF then reads 12 bytes per cell instead of B and v in FP64 (48 bytes), and
// in 2D forms each cell's once instead of once per corner. The FP64 form does
// this in 3D only (the EMF reads 24 bytes instead of 48); its cross product is
// the one the EMF would form from the same W, so results are unchanged.
template <int D, bool MHD, bool FINAL, int AX> constexpr bool eref_in_sweep() {
    return D >= 2 && AX == D - 1 && MHD && !FINAL && (sizeof(FT) == 4 || D == 3);
}
// From the cell state the sweep already holds in FP32.  A cell rebuilt by
// two lanes (segment or wavefront overlaps) is written twice, with the same
// bits.
template <class R>
__device__ __forceinline__ void store_eref(const SweepArgs& a, bool live, const int* idx, const P<R>& w) {
    if (!live || !a.eref[0]) return;
    const long e = a.fe.at(idx[0], idx[1], idx[2]);
    a.eref[0][e] = ET(w.By * w.vz - w.Bz * w.vy);  // cross(B, v), Arithmetic.cpp:359-365
    a.eref[1][e] = ET(w.Bz * w.vx - w.Bx * w.vz);
    a.eref[2][e] = ET(w.Bx * w.vy - w.By * w.vx);
}

// Conversions of the CTU-corrected half states.  The FP64 build keeps the
// reference's divisions.  The mixed build narrows these states to FP32 right
// after, so it divides through a refined reciprocal (v_rcp_f64 and two
// Newton steps, within an ulp of the quotient) in place of the correctly
// rounded FP64 division sequence; the energy difference of the pressure
// recovery is still formed in FP64.
#if DG_PREC == 32
__device__ __forceinline__ double rcp64(double x) {
    double r = __builtin_amdgcn_rcp(x);
    r = fma(fma(-x, r, 1.0), r, r);
    return fma(fma(-x, r, 1.0), r, r);
}
template <bool MHD> __device__ __forceinline__ U<double> cons_c(const P<double>& w, const Phys& ph) {
    U<double> u;
    u.rho = w.rho;
    u.mx = w.vx * w.rho; u.my = w.vy * w.rho; u.mz = w.vz * w.rho;
    u.E = w.p * (0.5 * ph._2_Gm1) + (w.rho / 2) * dot3(w.vx, w.vy, w.vz, w.vx, w.vy, w.vz);
    if constexpr (MHD) u.E += dot3(w.Bx, w.By, w.Bz, w.Bx, w.By, w.Bz) * ph.k1_8pi;
    if constexpr (MHD) { u.Bx = w.Bx; u.By = w.By; u.Bz = w.Bz; } else { u.Bx = u.By = u.Bz = 0; }
    return u;
}
template <bool MHD> __device__ __forceinline__ P<double> prim_c(const U<double>& u, const Phys& ph) {
    const double inv = rcp64(u.rho);
    P<double> w;
    w.rho = u.rho;
    w.vx = u.mx * inv; w.vy = u.my * inv; w.vz = u.mz * inv;
    double Et = u.E - dot3(u.mx, u.my, u.mz, u.mx, u.my, u.mz) * (0.5 * inv);
    if constexpr (MHD) Et -= dot3(u.Bx, u.By, u.Bz, u.Bx, u.By, u.Bz) * ph.k1_8pi;
    w.p = ph.gm1 * Et;
    if constexpr (MHD) { w.Bx = u.Bx; w.By = u.By; w.Bz = u.Bz; } else { w.Bx = w.By = w.Bz = 0; }
    return w;
}
#else
template <bool MHD> __device__ __forceinline__ U<double> cons_c(const P<double>& w, const Phys& ph) { return cons<MHD>(w, ph); }
template <bool MHD> __device__ __forceinline__ P<double> prim_c(const U<double>& u, const Phys& ph) { return prim<MHD>(u, ph); }
#endif

// CTU correction of the half states of one cell (double precision).
template <int D, bool MHD, int AX>
__device__ __forceinline__ void ctu_correct(const SweepArgs& a, long m, long q, double dBx, double dBy, double dBz,
                                            P<double>& L, P<double>& Rr) {
    const Ext& ce = a.ce;
    const Phys& ph = a.ph;
    auto fd = [&](int ax) { return usub<MHD>(load_u(a.Fp[ax], m, MHD), load_u(a.Fp[ax], m + ce.s[ax], MHD)); };
    if constexpr (!MHD) {  // correctState (hydro 2D): minus the transverse flux difference
        constexpr int T = AX == 0 ? 1 : 0;
        U<double> t = umul<false>(usub<false>(load_u(a.Fp[T], m + ce.s[T], false), load_u(a.Fp[T], m, false)), 0.5 * a.dt_d[T]);
        L = prim_c<false>(usub<false>(cons_c<false>(L, ph), t), ph);
        Rr = prim_c<false>(usub<false>(cons_c<false>(Rr, ph), t), ph);
        return;
    } else {
        const Ext& fe = a.fe;
        const long fx = fe.s[0], fy = fe.s[1], fz = fe.s[2];
        const double k8 = ph.k1_8pi;
        const P<double> Wc = load_p<double>(a.W, m, true);
        const double dt_dx = a.dt_d[0], dt_dy = a.dt_d[1], dt_dz = a.dt_d[2];
        const ET *E0x = a.E0[0], *E0y = a.E0[1], *E0z = a.E0[2];
        U<double> corr;
        if (D == 2) {
            if (AX == 0) {
                corr = umul<true>(umul<true>(fd(1), 0.5), dt_dy);
                corr.Bx = corr.By = corr.Bz = 0;
                double lim_z = minmod(-dBy, dBx);
                corr.E += k8 * (Wc.Bz * Wc.vz * lim_z);
                corr.Bz = 0.5 * dt_dy * (ed(E0x, q + fy) - ed(E0x, q)) + 0.5 * Wc.vz * lim_z;
            } else {
                corr = umul<true>(umul<true>(fd(0), 0.5), dt_dx);
                corr.Bx = corr.By = corr.Bz = 0;
                double lim_z = minmod(-dBx, dBy);
                corr.E += k8 * (Wc.Bz * Wc.vz * lim_z);
                corr.Bz = -0.5 * dt_dx * (ed(E0y, q + fx) - ed(E0y, q)) + 0.5 * Wc.vz * lim_z;
            }
        } else if (AX == 0) {
            corr = umul<true>(uadd<true>(umul<true>(fd(1), dt_dy), umul<true>(fd(2), dt_dz)), 0.5);
            corr.Bx = corr.By = corr.Bz = 0;
            double lim_y = minmod(-dBz, dBx), lim_z = minmod(-dBy, dBx);
            corr.E += k8 * (Wc.By * Wc.vy * lim_y + Wc.Bz * Wc.vz * lim_z);
            corr.By = -0.25 * dt_dz * (ed(E0x, q + fz) - ed(E0x, q) + ed(E0x, q + fy + fz) - ed(E0x, q + fy)) + 0.5 * Wc.vy * lim_y;
            corr.Bz = 0.25 * dt_dy * (ed(E0x, q + fy) - ed(E0x, q) + ed(E0x, q + fy + fz) - ed(E0x, q + fz)) + 0.5 * Wc.vz * lim_z;
        } else if (AX == 1) {
            corr = umul<true>(uadd<true>(umul<true>(fd(0), dt_dx), umul<true>(fd(2), dt_dz)), 0.5);
            corr.Bx = corr.By = corr.Bz = 0;
            double lim_x = minmod(-dBz, dBy), lim_z = minmod(-dBx, dBy);
            corr.E += k8 * (Wc.Bx * Wc.vx * lim_x + Wc.Bz * Wc.vz * lim_z);
            corr.Bx = 0.25 * dt_dz * (ed(E0y, q + fz) - ed(E0y, q) + ed(E0y, q + fx + fz) - ed(E0y, q + fx)) + 0.5 * Wc.vx * lim_x;
            corr.Bz = -0.25 * dt_dx * (ed(E0y, q + fx) - ed(E0y, q) + ed(E0y, q + fx + fz) - ed(E0y, q + fz)) + 0.5 * Wc.vz * lim_z;
        } else {
            corr = umul<true>(uadd<true>(umul<true>(fd(0), dt_dx), umul<true>(fd(1), dt_dy)), 0.5);
            corr.Bx = corr.By = corr.Bz = 0;
            double lim_x = minmod(-dBy, dBz), lim_y = minmod(-dBx, dBz);
            corr.E += k8 * (Wc.Bx * Wc.vx * lim_x + Wc.By * Wc.vy * lim_y);
            corr.Bx = -0.25 * dt_dy * (ed(E0z, q + fy) - ed(E0z, q) + ed(E0z, q + fx + fy) - ed(E0z, q + fx)) + 0.5 * Wc.vx * lim_x;
            corr.By = 0.25 * dt_dx * (ed(E0z, q + fx) - ed(E0z, q) + ed(E0z, q + fx + fy) - ed(E0z, q + fy)) + 0.5 * Wc.vy * lim_y;
        }
        U<double> uL = uadd<true>(cons_c<true>(L, ph), corr), uR = uadd<true>(cons_c<true>(Rr, ph), corr);
        const ET* Bh = a.Bh[AX];
        const long e = fe.s[AX];
        if (AX == 0) { uL.Bx = ed(Bh, q); uR.Bx = ed(Bh, q + e); }
        else if (AX == 1) { uL.By = ed(Bh, q); uR.By = ed(Bh, q + e); }
        else { uL.Bz = ed(Bh, q); uR.Bz = ed(Bh, q + e); }
        L = prim_c<true>(uL, ph);
        Rr = prim_c<true>(uR, ph);
    }
}

template <class R> __device__ __forceinline__ P<R> p_cast(const P<double>& w) {
    return P<R>{R(w.rho), R(w.p), R(w.vx), R(w.vy), R(w.vz), R(w.Bx), R(w.By), R(w.Bz)};
}
template <class R> __device__ __forceinline__ P<double> p_wide(const P<R>& w) {
    return P<double>{double(w.rho), double(w.p), double(w.vx), double(w.vy), double(w.vz), double(w.Bx), double(w.By), double(w.Bz)};
}

// Half states of cell (i,j,k) along AX, in the global frame, from the cell
// and its two neighbours along AX (wl, wr are read only for cells that are
// not on the outer ghost layer).
template <int D, bool MHD, class R, int AX, bool FINAL>
__device__ __forceinline__ void cell_states_from(const SweepArgs& a, int i, int j, int k, const P<R>& wl, const P<R>& wc,
                                                 const P<R>& wr, P<R>& L, P<R>& Rr) {
    const Ext& ce = a.ce;
    const Ext& fe = a.fe;
    const long m = ce.at(i, j, k);
    const int idx[3] = {i, j, k};
    const int g = ce.g;
    double bx = 0, by = 0, bz = 0;
    long q = 0;
    if (MHD && D >= 2) {
        q = fe.at(i, j, k);
        bx = (ld32(a.B[0], q + fe.s[0]) - ld32(a.B[0], q)) * a.dt_d[0];
        by = (ld32(a.B[1], q + fe.s[1]) - ld32(a.B[1], q)) * a.dt_d[1];
        if (D == 3) bz = (ld32(a.B[2], q + fe.s[2]) - ld32(a.B[2], q)) * a.dt_d[2];
    }
    if (idx[AX] >= -g + 1 && idx[AX] < ce.n[AX] + g - 1) {
        if (AX == 0) {
            muscl<MHD, R>(wl, L, wc, Rr, wr, R(a.dt_d[0]), R(bx), R(by), R(bz), a.ph);
        } else if (AX == 1) {
            P<R> l, r;
            muscl<MHD, R>(swapXY(wl), l, swapXY(wc), r, swapXY(wr), R(a.dt_d[1]), R(by), R(bx), R(bz), a.ph);
            L = swapXY(l); Rr = swapXY(r);
        } else {
            P<R> l, r;
            muscl<MHD, R>(swapXZ(wl), l, swapXZ(wc), r, swapXZ(wr), R(a.dt_d[2]), R(bz), R(by), R(bx), a.ph);
            L = swapXZ(l); Rr = swapXZ(r);
        }
    } else {
        L = wc; Rr = wc;
    }
    if (MHD && D >= 2) {  // copyFaceFields: normal B := face B
        const double* Bn = a.B[AX];
        const R b0 = R(ld32(Bn, q)), b1 = R(ld32(Bn, q + fe.s[AX]));
        if (AX == 0) { L.Bx = b0; Rr.Bx = b1; }
        else if (AX == 1) { L.By = b0; Rr.By = b1; }
        else { L.Bz = b0; Rr.Bz = b1; }
    }
    if constexpr (FINAL) {
        P<double> l = p_wide(L), r = p_wide(Rr);
        ctu_correct<D, MHD, AX>(a, m, q, bx, by, bz, l, r);
        L = p_cast<R>(l);
        Rr = p_cast<R>(r);
    }
}

// The same, reading the three cells.
template <int D, bool MHD, class R, int AX, bool FINAL>
__device__ __forceinline__ void cell_states(const SweepArgs& a, int i, int j, int k, P<R>& L, P<R>& Rr, P<R>* wc_out = nullptr) {
    const long m = a.ce.at(i, j, k);
    const int c = AX == 0 ? i : AX == 1 ? j : k;
    const P<R> wc = load_p<R>(a.W, m, MHD);
    if (wc_out) *wc_out = wc;
    if (c >= -a.ce.g + 1 && c < a.ce.n[AX] + a.ce.g - 1) {
        const long off = a.ce.s[AX];
        const P<R> wl = load_p<R>(a.W, m - off, MHD), wr = load_p<R>(a.W, m + off, MHD);
        cell_states_from<D, MHD, R, AX, FINAL>(a, i, j, k, wl, wc, wr, L, Rr);
    } else {
        cell_states_from<D, MHD, R, AX, FINAL>(a, i, j, k, wc, wc, wc, L, Rr);
    }
}

template <bool MHD, class R, int AX>
__device__ __forceinline__ void face_flux(const SweepArgs& a, P<R> l, P<R> r, long m) {
    if (AX == 1) { l = swapXY(l); r = swapXY(r); }
    if (AX == 2) { l = swapXZ(l); r = swapXZ(r); }
    int fail = RF_OK;
    U<R> f = riemann_x<MHD, R>(l, r, R(a.dt_d[AX]), a.ph, &fail);
    if (fail) report(a.err, a.stage, (unsigned long long)m, fail == RF_BAD_INPUT ? E_RIEMANN_INPUT : E_RIEMANN_SOLUTION);
    if (AX == 1) f = swapXY(f);
    if (AX == 2) f = swapXZ(f);
    store_u<R>(a.Fo, m, f, MHD);
}

// Windowed walkers (stencil cells carried in registers) for the FP32 form;
// the FP64 walkers are register-bound (the window takes them to 256 VGPRs,
// one wave per SIMD): they read the cell ahead early and re-read the two
// behind it.  The windowed FP64 instantiation exists as a check of the
// window's cell bookkeeping: its results must equal the oracle's bit for bit.
template <class R> constexpr bool march_window() { return sizeof(R) == 4; }

// Two values of the walker's precision, one LDS access.
template <class R> struct Pair2;
template <> struct Pair2<float> { using T = float2; };
template <> struct Pair2<double> { using T = double2; };

// Strided axis: one thread walks `chunk` consecutive faces.  PARKALL: the
// windowed walker parks its carried state in LDS in the predictor too (the
// batched walkers); without it, a must be the kernel's own argument (the
// single-grid walkers).
template <int D, bool MHD, class R, int AX, bool FINAL, bool WIN = march_window<R>(), bool PARKALL = false>
__device__ __forceinline__ void sweep_march_body(const SweepArgs& a, long t) {
    constexpr int T0 = AX == 0 ? 1 : 0;            // slower transverse dim (3D only)
    constexpr int T1 = D == 3 ? (AX == 2 ? 1 : 2) : (AX == 0 ? 1 : 0);  // fastest transverse dim
    const long n1 = a.hi[T1] - a.lo[T1];
    const long n0 = D == 3 ? a.hi[T0] - a.lo[T0] : 1;
    const int len = a.hi[AX] - a.lo[AX];
    const int nch = (len + a.chunk - 1) / a.chunk;
    if (t >= (long)nch * n0 * n1) return;
    int idx[3] = {0, 0, 0};
    idx[T1] = a.lo[T1] + (int)(t % n1);
    long r = t / n1;
    if (D == 3) { idx[T0] = a.lo[T0] + (int)(r % n0); r /= n0; }
    const int f0 = a.lo[AX] + (int)r * a.chunk;
    const int f1 = min(a.hi[AX], f0 + a.chunk);
    P<R> Rp = pzero<R>(), Lc, Rc;
    idx[AX] = f0 - 1;
    // One loop body rebuilds every cell's states, the walker's first cell (f0 -
    // 1, no face of its own) included: with FMA contraction a second inlined
    // copy may round differently, and a periodic image face rebuilt by the
    // other copy would then take a different update than its twin.
    const long s = a.ce.s[AX];
    const int clo = -a.ce.g, chi = a.ce.n[AX] + a.ce.g;  // cells that exist along AX
    long m = a.ce.at(idx[0], idx[1], idx[2]);
    auto ld = [&](int c, long mm) { return c >= clo && c < chi ? load_p<double>(a.W, mm, MHD) : pzero<double>(); };
    if constexpr (WIN) {
        // The three cells of the MUSCL stencil are carried as the walker
        // advances; the cell two ahead is read (as FP64) before the Riemann
        // solve of the current face and narrowed when it is used.  The
        // predictor keeps the two cells behind it and the right state of the
        // last cell in registers.  The corrector, which also carries the CTU
        // corrections' operands, needs 159 registers (3 waves per SIMD) that
        // way; where the walker's step is short (not the 3D x walker, whose
        // step is a whole plane) it parks them in LDS instead, one slot per
        // thread, and so do both batched walkers.  The parked single-grid
        // corrector also reads its arguments again from the kernel-argument
        // segment at each face, through an opaque copy of its address, so
        // that no value derived from them (nor a constant the compiler would
        // hoist with them) stays in a register across the loop: 116-118
        // registers, 4 waves, where it would otherwise spill at 4.  (At 4
        // waves the 3D x corrector ran 3% faster or 5% slower from run to
        // run, the y corrector 8% faster in every run.)
        constexpr bool PARK = (FINAL && !(D == 3 && AX == 0)) || PARKALL;
        constexpr bool RELOAD = PARK && !PARKALL;
        using R2 = typename Pair2<R>::T;
        __shared__ R2 park[3 * 4 * 256];  // [slot][field pair][thread]: 256 threads per workgroup
        R2* pk = park + threadIdx.x;
        auto put = [&](int slot, const P<R>& w) {
            R2* q = pk + slot * 1024;
            q[0] = R2{w.rho, w.p}; q[256] = R2{w.vx, w.vy}; q[512] = R2{w.vz, w.Bx}; q[768] = R2{w.By, w.Bz};
        };
        auto get = [&](int slot) {
            const R2* q = pk + slot * 1024;
            const R2 a0 = q[0], a1 = q[256], a2 = q[512], a3 = q[768];
            return P<R>{a0.x, a0.y, a1.x, a1.y, a2.x, a2.y, a3.x, a3.y};
        };
        P<R> wl = pzero<R>(), wc = pzero<R>();
        if constexpr (PARK) {
            put(0, p_cast<R>(ld(f0 - 2, m - s)));
            put(1, p_cast<R>(ld(f0 - 1, m)));
        } else {
            wl = p_cast<R>(ld(f0 - 2, m - s));
            wc = p_cast<R>(ld(f0 - 1, m));
        }
        P<double> nxt = ld(f0, m + s);
        int sl = 0;  // parked: slot of the stencil's left cell; its centre cell is in slot 1 - sl
        typedef const __attribute__((address_space(4))) SweepArgs* KArg;
        KArg kp = (KArg)__builtin_amdgcn_kernarg_segment_ptr();
        auto step = [&](const SweepArgs& a, int f) {
            idx[AX] = f;
            const P<R> wr = p_cast<R>(nxt);
            if constexpr (PARK) {
                wl = get(sl);
                wc = get(1 - sl);
                put(sl, wr);  // the left cell's slot takes the new leading cell
            }
            cell_states_from<D, MHD, R, AX, FINAL>(a, idx[0], idx[1], idx[2], wl, wc, wr, Lc, Rc);
            nxt = f + 2 >= clo && f + 2 < chi ? load_p<double>(a.W, m + 2 * s, MHD) : pzero<double>();
            if constexpr (PARK) {
                if (f >= f0) Rp = get(2);
                put(2, Rc);
                // Rc's only copy for the next face is the one in LDS
                asm volatile("" ::: "memory");
                if (f >= f0) face_flux<MHD, R, AX>(a, Rp, Lc, m);  // uniform per wavefront
                sl = 1 - sl;
            } else {
                if (f >= f0) face_flux<MHD, R, AX>(a, Rp, Lc, m);
                Rp = Rc;
                wl = wc;
                wc = wr;
            }
            m += s;
        };
        for (int f = f0 - 1; f < f1; f++) {
            if constexpr (RELOAD) {
                asm volatile("" : "+s"(kp));
                step(*(const SweepArgs*)kp, f);
            } else {
                step(a, f);
            }
        }
    } else {
        // The cell ahead of the stencil is read before the Riemann solve of
        // the current face; the two behind it are read again (cache hits):
        // holding them too takes the FP64 walker to 256 registers.
        P<double> nxt = ld(f0, m + s);
        for (int f = f0 - 1; f < f1; f++) {
            idx[AX] = f;
            cell_states_from<D, MHD, R, AX, FINAL>(a, idx[0], idx[1], idx[2], p_cast<R>(ld(f - 1, m - s)), p_cast<R>(ld(f, m)),
                                                   p_cast<R>(nxt), Lc, Rc);
            nxt = ld(f + 2, m + 2 * s);
            if (f >= f0) face_flux<MHD, R, AX>(a, Rp, Lc, m);
            Rp = Rc;
            m += s;
        }
    }
}
// Occupancy floor of the FP32 walkers: four waves per SIMD for the predictor
// (133 registers, 8 bytes per lane spilled at four, 5-8% faster than three)
// and the parked corrector (116-118 registers).  1 = the compiler's choice.
template <int D, int AX, class R, bool FINAL> constexpr int march_min_waves() {
    return DG_SWEEP_WAVES == 0 && sizeof(R) == 4 && (!FINAL || !(D == 3 && AX == 0)) ? 4 : 1;
}
template <int D, bool MHD, class R, int AX, bool FINAL>
KS __attribute__((amdgpu_waves_per_eu(march_min_waves<D, AX, R, FINAL>()))) __global__ void k_sweep_march(SweepArgs a) {
    sweep_march_body<D, MHD, R, AX, FINAL>(a, tid_xcd());
}
template <int D, bool MHD, class R, int AX, bool FINAL>
KS __global__ void k_sweep_march_window(SweepArgs a) { sweep_march_body<D, MHD, R, AX, FINAL, true>(a, tid()); }

template <class R> __device__ __forceinline__ P<R> shfl_up1(const P<R>& w) {
    return P<R>{__shfl_up(w.rho, 1), __shfl_up(w.p, 1), __shfl_up(w.vx, 1), __shfl_up(w.vy, 1),
                __shfl_up(w.vz, 1), __shfl_up(w.Bx, 1), __shfl_up(w.By, 1), __shfl_up(w.Bz, 1)};
}

// Contiguous axis in segments (the batched y sweep): lanes are grouped into
// segments of SW lanes (64, 32 or 16); lane l of a segment rebuilds cell
// c0-1+l and lanes 1..SW-1 take the left state from the lane below, so a
// segment computes SW-1 faces.  With SW = 64 the segment is the wavefront
// and its index is wave-uniform.
template <int D, bool MHD, class R, int AX, bool FINAL, int SW = 64>
__device__ __forceinline__ void sweep_wave_body(const SweepArgs& a, long seg_id, int lane) {
    constexpr int T0 = 0;                 // slowest transverse dim
    constexpr int T1 = D == 3 ? 1 : 0;    // faster transverse dim
    const int ll = lane % SW;             // lane within the segment
    const int len = a.hi[AX] - a.lo[AX];
    const int nseg = (len + SW - 2) / (SW - 1);
    const int n1 = D == 3 ? a.hi[T1] - a.lo[T1] : 1;
    const int n0 = a.hi[T0] - a.lo[T0];
    if (seg_id >= (long)nseg * n0 * n1) return;  // uniform per segment
    int idx[3] = {0, 0, 0};
    int r = (int)seg_id;  // < 2^31 segments: 32-bit division (the host sizes grids far below)
    const int seg = r % nseg;
    r /= nseg;
    if (D == 3) { idx[T1] = a.lo[T1] + r % n1; r /= n1; }
    idx[T0] = a.lo[T0] + r;
    const int c = a.lo[AX] + seg * (SW - 1) - 1 + ll;
    const bool live = c < a.hi[AX];
    P<R> L = pzero<R>(), Rr = pzero<R>(), wc = pzero<R>();
    idx[AX] = c;
    if (live) cell_states<D, MHD, R, AX, FINAL>(a, idx[0], idx[1], idx[2], L, Rr, &wc);
    if constexpr (eref_in_sweep<D, MHD, FINAL, AX>()) store_eref(a, live, idx, wc);
    const P<R> Rl = shfl_up1(Rr);  // lane 0 of a segment takes the lane below's value and leaves it unused
    if (ll >= 1 && live) face_flux<MHD, R, AX>(a, Rl, L, a.ce.at(idx[0], idx[1], idx[2]));
}
// Contiguous axis, streamed: the cells to rebuild, lo-1 .. hi-1 of every
// line (line after line, which is also their order in memory when the range
// spans the row), are dealt to wavefronts 63 at a time, each wavefront
// taking one cell more than its 63 (the first cell of the next, rebuilt
// twice).  Lane l >= 1 forms the face between its cell and lane l-1's unless
// its cell is the first of a line.  Every face is formed once, and 63 of 64
// lanes form a face except at line starts, where fixed segments of 63 faces
// per line (9 wavefronts for the 517 faces of a 512^3 predictor line) left a
// tenth of the lanes idle.
template <int D, bool MHD, class R, int AX, bool FINAL>
__device__ __forceinline__ void sweep_stream_body(const SweepArgs& a, int w, int lane) {
    static_assert(AX == D - 1, "contiguous axis");
    const int len = a.hi[AX] - a.lo[AX];
    const int cells = len + 1;  // per line
    const int n1 = D == 3 ? a.hi[1] - a.lo[1] : 1;
    const int n0 = D >= 2 ? a.hi[0] - a.lo[0] : 1;
    const unsigned g = 63u * (unsigned)w + (unsigned)lane;  // < 2^31: the host sizes grids far below
    const bool live = g < (unsigned)(n0 * n1 * cells);
    const unsigned line = g / (unsigned)cells, c = g - line * (unsigned)cells;
    int idx[3] = {0, 0, 0};
    if (D == 3) { idx[0] = a.lo[0] + (int)(line / (unsigned)n1); idx[1] = a.lo[1] + (int)(line % (unsigned)n1); }
    if (D == 2) idx[0] = a.lo[0] + (int)line;
    idx[AX] = a.lo[AX] - 1 + (int)c;
    P<R> L = pzero<R>(), Rr = pzero<R>(), wc = pzero<R>();
    if (live) cell_states<D, MHD, R, AX, FINAL>(a, idx[0], idx[1], idx[2], L, Rr, &wc);
    if constexpr (eref_in_sweep<D, MHD, FINAL, AX>()) store_eref(a, live, idx, wc);
    const P<R> Rl = shfl_up1(Rr);  // lane 0 takes the lane below's value and leaves it unused
    if (lane >= 1 && live && c >= 1) face_flux<MHD, R, AX>(a, Rl, L, a.ce.at(idx[0], idx[1], idx[2]));
}
// Occupancy floor of the contiguous-axis sweeps: the FP64 ones need 176
// registers, 8 short of three waves per SIMD; at three they spill 12 bytes
// per lane.  1 = the compiler's choice.
template <class R> constexpr int wave_min_waves() { return DG_SWEEP_WAVES == 0 && sizeof(R) == 8 ? 3 : 1; }
template <int D, bool MHD, class R, int AX, bool FINAL>
KS __attribute__((amdgpu_waves_per_eu(wave_min_waves<R>()))) __global__ void k_sweep_wave(SweepArgs a) {
    // the wavefront index is uniform: keep it in a scalar register
    const int w = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    sweep_stream_body<D, MHD, R, AX, FINAL>(a, (int)block_xcd() * (int)(blockDim.x >> 6) + w, threadIdx.x & 63);
}

// ===================================================== CTU hydro (2D/3D)
// correctState (CTU.cpp:264-301): states along SAX corrected by the flux
// difference of FAX.
template <int D, class R>
KB __global__ void k_correct(Soa HL, Soa HR, FSoa F, Ext ce, Box box, int fax, double dt_dL, Phys ph) {
    int i, j, k;
    if (!box.point(tid(), i, j, k)) return;
    const long m = ce.at(i, j, k);
    U<double> t = umul<false>(usub<false>(load_u(F, m + ce.s[fax], false), load_u(F, m, false)), dt_dL);
    P<double> l = load_p<double>(HL, m, false), r = load_p<double>(HR, m, false);
    store_p<double>(HL, m, prim<false>(usub<false>(cons<false>(l, ph), t), ph), false);
    store_p<double>(HR, m, prim<false>(usub<false>(cons<false>(r, ph), t), ph), false);
}

// computeCTUFlux_X/Y/Z (CTU.cpp:306-370): a flux between states corrected by
// the transverse flux T (3D hydro, 12-solve CTU).
template <class R, int AX>
KB __global__ void k_ctu_flux(Soa HL, Soa HR, FSoa T, FSoa F, Ext ce, Box box, int tax, double dt_d, double dt_t, Phys ph,
                              unsigned long long* err, int stage) {
    int i, j, k;
    if (!box.point(tid(), i, j, k)) return;
    const long m = ce.at(i, j, k), e = ce.s[AX], te = ce.s[tax];
    auto corr = [&](const Soa& S, long q) {
        U<double> t = umul<false>(usub<false>(load_u(T, q + te, false), load_u(T, q, false)), dt_t);
        return prim<false>(usub<false>(cons<false>(load_p<double>(S, q, false), ph), t), ph);
    };
    P<double> l = corr(HR, m - e), r = corr(HL, m);
    P<R> lr = {R(l.rho), R(l.p), R(l.vx), R(l.vy), R(l.vz), 0, 0, 0};
    P<R> rr = {R(r.rho), R(r.p), R(r.vx), R(r.vy), R(r.vz), 0, 0, 0};
    if (AX == 1) { lr = swapXY(lr); rr = swapXY(rr); }
    if (AX == 2) { lr = swapXZ(lr); rr = swapXZ(rr); }
    int fail = RF_OK;
    U<R> f = riemann_x<false, R>(lr, rr, R(dt_d), ph, &fail);
    if (fail) {
        // the reference walks these loops with the flux axis innermost
        const int ia[3] = {i + ce.g, j + ce.g, k + ce.g};
        const int o0 = AX == 0 ? 1 : 0, o1 = AX == 2 ? 1 : 2;
        unsigned long long seq = ((unsigned long long)ia[o0] * 65536ull + (unsigned long long)ia[o1]) * 65536ull + ia[AX];
        report(err, stage, seq, fail == RF_BAD_INPUT ? E_RIEMANN_INPUT : E_RIEMANN_SOLUTION);
    }
    if (AX == 1) f = swapXY(f);
    if (AX == 2) f = swapXZ(f);
    store_u<R>(F, m, f, false);
}

// ================================================================ MHD CT
__device__ __forceinline__ double upwind_corr(double m, double fm, double fp, double rm, double rp) {  // Electric.cpp:96-101
    if (m > 1e-18) return 0.25 * (fm - rm);
    if (m < -1e-18) return 0.25 * (fp - rp);
    return 0.125 * (fm + fp - rm - rp);
}
// Flux SoA field indices (U order)
enum { U_RHO = 0, U_BX = 5, U_BY = 6, U_BZ = 7 };

// Reference EMF of a cell: cross(w.B, w.v) (Electric.cpp:108-125); `Ws`
// selects whether the cell state comes from W (predictor) or from the
// half-step Eref arrays (corrector).
struct ERef {
    const ET *x, *y, *z;
};

// 2D corner EMF (Electric.cpp:24-35, 52-71) at edge (i,j).  With `fromW`,
// Eref is computed from the cell states (predictor); otherwise it is read.
template <bool FROM_W>
__device__ __forceinline__ void emf2_body(const FSoa& FX, const FSoa& FY, const Soa& W, const ERef& ref, const Ext& ce, const Ext& fe,
                                          const Box& box, int upwind, ET* Ex, ET* Ey, ET* Ez, long t) {
    int i, j, k;
    if (!box.point(t, i, j, k)) return;
    auto f = [&](const FSoa& F, int fld, int a, int b) { return double(F(fld, ce.at(a, b, 0))); };
    double ex = -f(FY, U_BZ, i, j);
    double ey = f(FX, U_BZ, i, j);
    double ez = (f(FY, U_BX, i - 1, j) - f(FX, U_BY, i, j - 1) + f(FY, U_BX, i, j) - f(FX, U_BY, i, j)) * 0.25;
    if (upwind) {
        double rx[2][2], ry[2][2], rz[2][2];  // [di][dj] for cells (i-1+di, j-1+dj)
        for (int di = 0; di < 2; di++)
            for (int dj = 0; dj < 2; dj++) {
                if (FROM_W) {
                    const long c = ce.at(i - 1 + di, j - 1 + dj, 0);
                    cross(W(F_BX, c), W(F_BY, c), W(F_BZ, c), W(F_VX, c), W(F_VY, c), W(F_VZ, c), rx[di][dj], ry[di][dj], rz[di][dj]);
                } else {
                    const long c = fe.at(i - 1 + di, j - 1 + dj, 0);
                    rx[di][dj] = ed(ref.x, c); ry[di][dj] = ed(ref.y, c); rz[di][dj] = ed(ref.z, c);
                }
            }
        ez += upwind_corr(f(FX, U_RHO, i, j), f(FY, U_BX, i - 1, j), f(FY, U_BX, i, j), rz[0][1], rz[1][1]);
        ez += upwind_corr(f(FX, U_RHO, i, j - 1), f(FY, U_BX, i - 1, j), f(FY, U_BX, i, j), rz[0][0], rz[1][0]);
        ez += upwind_corr(f(FY, U_RHO, i, j), -f(FX, U_BY, i, j - 1), -f(FX, U_BY, i, j), rz[1][0], rz[1][1]);
        ez += upwind_corr(f(FY, U_RHO, i - 1, j), -f(FX, U_BY, i, j - 1), -f(FX, U_BY, i, j), rz[0][0], rz[0][1]);
        ex += 0.25 * (-2 * f(FY, U_BZ, i, j) - rx[1][1] - rx[1][0]);
        ey += 0.25 * (2 * f(FX, U_BZ, i, j) - ry[1][1] - ry[0][1]);
    }
    const long e = fe.at(i, j, 0);
    Ex[e] = ex; Ey[e] = ey; Ez[e] = ez;
}
template <bool FROM_W>
KB __global__ void k_emf2(FSoa FX, FSoa FY, Soa W, ERef ref, Ext ce, Ext fe, Box box, int upwind, ET* Ex, ET* Ey,
                          ET* Ez) {
    emf2_body<FROM_W>(FX, FY, W, ref, ce, fe, box, upwind, Ex, Ey, Ez, tid());
}

// 3D corner EMF (Electric.cpp:38-47, 74-100) of the edge triple at (i,j,k).
// fx/fy/fz(f, di, dj, dk) give flux component f (0 mass, 1 and 2 the two
// transverse B components in x,y,z order) of the x/y/z flux at (i+di, j+dj,
// k+dk), d in {-1, 0}; R(c, di, dj, dk) component c of the reference EMF of
// that cell.
template <class FA, class GA, class HA, class RA>
__device__ __forceinline__ void emf3_eval(FA fx, GA fy, HA fz, RA R, int upwind, double& ex, double& ey, double& ez) {
    // fx: {rho, By, Bz}; fy: {rho, Bx, Bz}; fz: {rho, Bx, By}
    ex = (fz(2, 0, -1, 0) - fy(2, 0, 0, -1) + fz(2, 0, 0, 0) - fy(2, 0, 0, 0)) * 0.25;
    ey = (fx(2, 0, 0, -1) - fz(1, -1, 0, 0) + fx(2, 0, 0, 0) - fz(1, 0, 0, 0)) * 0.25;
    ez = (fy(1, -1, 0, 0) - fx(1, 0, -1, 0) + fy(1, 0, 0, 0) - fx(1, 0, 0, 0)) * 0.25;
    if (upwind) {
        ey += upwind_corr(fx(0, 0, 0, 0), -fz(1, -1, 0, 0), -fz(1, 0, 0, 0), R(1, -1, 0, 0), R(1, 0, 0, 0));
        ey += upwind_corr(fx(0, 0, 0, -1), -fz(1, -1, 0, 0), -fz(1, 0, 0, 0), R(1, -1, 0, -1), R(1, 0, 0, -1));
        ez += upwind_corr(fx(0, 0, 0, 0), fy(1, -1, 0, 0), fy(1, 0, 0, 0), R(2, -1, 0, 0), R(2, 0, 0, 0));
        ez += upwind_corr(fx(0, 0, -1, 0), fy(1, -1, 0, 0), fy(1, 0, 0, 0), R(2, -1, -1, 0), R(2, 0, -1, 0));
        ex += upwind_corr(fy(0, 0, 0, 0), fz(2, 0, -1, 0), fz(2, 0, 0, 0), R(0, 0, -1, 0), R(0, 0, 0, 0));
        ex += upwind_corr(fy(0, 0, 0, -1), fz(2, 0, -1, 0), fz(2, 0, 0, 0), R(0, 0, -1, -1), R(0, 0, 0, -1));
        ez += upwind_corr(fy(0, 0, 0, 0), -fx(1, 0, -1, 0), -fx(1, 0, 0, 0), R(2, 0, -1, 0), R(2, 0, 0, 0));
        ez += upwind_corr(fy(0, -1, 0, 0), -fx(1, 0, -1, 0), -fx(1, 0, 0, 0), R(2, -1, -1, 0), R(2, -1, 0, 0));
        ex += upwind_corr(fz(0, 0, 0, 0), -fy(2, 0, 0, -1), -fy(2, 0, 0, 0), R(0, 0, 0, -1), R(0, 0, 0, 0));
        ex += upwind_corr(fz(0, 0, -1, 0), -fy(2, 0, 0, -1), -fy(2, 0, 0, 0), R(0, 0, -1, -1), R(0, 0, -1, 0));
        ey += upwind_corr(fz(0, 0, 0, 0), fx(2, 0, 0, -1), fx(2, 0, 0, 0), R(1, 0, 0, -1), R(1, 0, 0, 0));
        ey += upwind_corr(fz(0, -1, 0, 0), fx(2, 0, 0, -1), fx(2, 0, 0, 0), R(1, -1, 0, -1), R(1, -1, 0, 0));
    }
}
// flux component slots of emf3_eval per axis: x {rho, By, Bz}, y {rho, Bx, Bz}, z {rho, Bx, By}
__device__ __forceinline__ int emf_fld(int ax, int f) {
    return f == 0 ? U_RHO : ax == 0 ? (f == 1 ? U_BY : U_BZ) : ax == 1 ? (f == 1 ? U_BX : U_BZ) : (f == 1 ? U_BX : U_BY);
}

// One thread per edge triple, every operand read from memory.
template <bool FROM_W>
KB __global__ void k_emf3(FSoa FX, FSoa FY, FSoa FZ, Soa W, ERef ref, Ext ce, Ext fe, Box box, int upwind, ET* Ex,
                          ET* Ey, ET* Ez) {
    int i, j, k;
    if (!box.point(tid(), i, j, k)) return;
    auto fx = [&](int f, int a, int b, int c) { return double(FX(emf_fld(0, f), ce.at(i + a, j + b, k + c))); };
    auto fy = [&](int f, int a, int b, int c) { return double(FY(emf_fld(1, f), ce.at(i + a, j + b, k + c))); };
    auto fz = [&](int f, int a, int b, int c) { return double(FZ(emf_fld(2, f), ce.at(i + a, j + b, k + c))); };
    auto R = [&](int comp, int a, int b, int c) -> double {
        if (FROM_W) {
            const long q = ce.at(i + a, j + b, k + c);
            double x, y, z;
            cross(W(F_BX, q), W(F_BY, q), W(F_BZ, q), W(F_VX, q), W(F_VY, q), W(F_VZ, q), x, y, z);
            return comp == 0 ? x : comp == 1 ? y : z;
        }
        const long q = fe.at(i + a, j + b, k + c);
        return comp == 0 ? ed(ref.x, q) : comp == 1 ? ed(ref.y, q) : ed(ref.z, q);
    };
    double ex, ey, ez;
    emf3_eval(fx, fy, fz, R, upwind, ex, ey, ez);
    const long e = fe.at(i, j, k);
    Ex[e] = ex; Ey[e] = ey; Ez[e] = ez;
}

// The same edges blocked for reuse: a workgroup owns a (j,k) tile of
// TJ x TK edges and walks a run of i-planes.  Each cell's operands (9 flux
// components and its stored reference EMF) are read from memory once per
// plane into LDS, with a one-cell halo on the low j and k sides; the previous
// plane stays in the other LDS buffer for the i-1 neighbours.  Every edge is
// evaluated by emf3_eval, as in k_emf3.
//
// The tile and its halo are one cell per thread (3 x 63 edges, 4 x 64
// cells), and each thread reads the next plane's cell into registers before
// the current plane's edges are evaluated, so the reads' latency overlaps
// that work.
struct EmfTile {
    static constexpr int TK = 63, TJ = 3, CELLS = (TJ + 1) * (TK + 1);
};
static_assert(EmfTile::CELLS == 256, "one halo-extended tile cell per thread");
// emf3_eval reads the plane below the edge (offset a = -1) only through the
// y and z fluxes' mass and first B slots and the reference EMF's y and z
// components; those are kept for two planes (ring slot = plane & 1), the
// rest (x fluxes, the y and z fluxes' second B slot, the x reference EMF)
// for the current plane only.  Half the LDS of two full planes: more
// workgroups per CU.
template <int CELLS> struct EmfPlanes {
    double rk[2][2][CELLS];  // reference EMF y, z; [ring][comp - 1][cell]
    FT fk[2][4][CELLS];      // FY slots 0, 1 and FZ slots 0, 1; [ring][..][cell]
    double rx[CELLS];        // reference EMF x, current plane
    FT fo[5][CELLS];         // FX slots 0, 1, 2, FY slot 2, FZ slot 2, current plane
};
// FP32 operands: five workgroups per CU fit in LDS (23-30 KB each), and at
// most 96 registers keep five waves per SIMD for them.
template <class T> constexpr int emf_tiled_waves() { return sizeof(T) == 4 ? 5 : 1; }
KB __attribute__((amdgpu_waves_per_eu(emf_tiled_waves<FT>()))) __global__ void k_emf3_tiled(
    FSoa FX, FSoa FY, FSoa FZ, ERef ref, Ext ce, Ext fe, Box box, int upwind, int iplanes, ET* Ex, ET* Ey, ET* Ez) {
    constexpr int TK = EmfTile::TK, TJ = EmfTile::TJ, CELLS = EmfTile::CELLS;
    __shared__ EmfPlanes<CELLS> S;
    const int nkt = (box.hi[2] - box.lo[2] + TK - 1) / TK;
    const int njt = (box.hi[1] - box.lo[1] + TJ - 1) / TJ, njq = (njt + 3) / 4;
    // runs of 4 j-neighbour tiles at one k, each run on one XCD (tid_xcd's
    // order), so that the low-j halo row a tile reads is mostly its
    // neighbour's own row in the same L2
    int b = (int)block_xcd();
    const int jq = b % 4;
    b /= 4;
    const int kt = b % nkt;
    b /= nkt;
    const int jt = (b % njq) * 4 + jq;
    const int ic = b / njq;
    if (jt >= njt) return;  // the whole workgroup
    const int k0 = box.lo[2] + kt * TK, j0 = box.lo[1] + jt * TJ;
    const int ib = box.lo[0] + ic * iplanes, ie = min(box.hi[0], ib + iplanes);
    // a cell c of a plane: read into registers (fetch), then into LDS
    // (stash): the kept values into ring slot i & 1, the others (when `own`)
    // into the current-plane arrays.  Cells outside the box are skipped: no
    // edge reads them.
    struct CellVals {
        FT f[9];      // FY slots 0, 1, FZ slots 0, 1, FX slots 0, 1, 2, FY slot 2, FZ slot 2
        double w[3];  // the reference EMF's x, y, z
    };
    auto in_box = [&](int c) { return j0 + c / (TK + 1) - 1 < box.hi[1] && k0 + c % (TK + 1) - 1 < box.hi[2]; };
    auto fetch = [&](int c, int i, bool own, CellVals& v) {
        const int cj = j0 + c / (TK + 1) - 1, ck = k0 + c % (TK + 1) - 1;
        const long q = ce.at(i, cj, ck);
        v.f[0] = FY(emf_fld(1, 0), q);
        v.f[1] = FY(emf_fld(1, 1), q);
        v.f[2] = FZ(emf_fld(2, 0), q);
        v.f[3] = FZ(emf_fld(2, 1), q);
        if (own) {
            for (int f = 0; f < 3; f++) v.f[4 + f] = FX(emf_fld(0, f), q);
            v.f[7] = FY(emf_fld(1, 2), q);
            v.f[8] = FZ(emf_fld(2, 2), q);
        }
        if (upwind) {
            const long e = fe.at(i, cj, ck);
            v.w[0] = ed(ref.x, e); v.w[1] = ed(ref.y, e); v.w[2] = ed(ref.z, e);
        }
    };
    auto stash = [&](int c, int i, bool own, const CellVals& v) {
        const int rs = i & 1;
        for (int f = 0; f < 4; f++) S.fk[rs][f][c] = v.f[f];
        if (own)
            for (int f = 0; f < 5; f++) S.fo[f][c] = v.f[4 + f];
        if (upwind) {
            S.rk[rs][0][c] = v.w[1];
            S.rk[rs][1][c] = v.w[2];
            if (own) S.rx[c] = v.w[0];
        }
    };
    const int tk = threadIdx.x % TK, tj = threadIdx.x / TK;
    const int j = j0 + tj, k = k0 + tk;
    const bool live = tj < TJ && j < box.hi[1] && k < box.hi[2];
    auto cell = [&](int dj, int dk) { return (tj + dj + 1) * (TK + 1) + (tk + dk + 1); };
    const int mc = threadIdx.x;  // this thread's cell
    const bool mine = in_box(mc);
    CellVals v;
    if (mine) { fetch(mc, ib - 1, false, v); stash(mc, ib - 1, false, v); }
    if (mine) fetch(mc, ib, true, v);
    for (int i = ib; i < ie; i++) {
        if (mine) stash(mc, i, true, v);
        __syncthreads();
        if (mine && i + 1 < ie) fetch(mc, i + 1, true, v);  // waited for at the next stash, after this plane's edges
        auto ring = [&](int a) { return (i + a) & 1; };  // a = 0 or -1
        // the x fluxes and the reference EMF's x component are read at a = 0 only (emf3_eval)
        auto fx = [&](int f, int, int b2, int c2) { return double(S.fo[f][cell(b2, c2)]); };
        auto fy = [&](int f, int a, int b2, int c2) { return double(f < 2 ? S.fk[ring(a)][f][cell(b2, c2)] : S.fo[3][cell(b2, c2)]); };
        auto fz = [&](int f, int a, int b2, int c2) { return double(f < 2 ? S.fk[ring(a)][2 + f][cell(b2, c2)] : S.fo[4][cell(b2, c2)]); };
        auto R = [&](int comp, int a, int b2, int c2) { return comp == 0 ? S.rx[cell(b2, c2)] : S.rk[ring(a)][comp - 1][cell(b2, c2)]; };
        if (live) {
            double ex, ey, ez;
            emf3_eval(fx, fy, fz, R, upwind, ex, ey, ez);
            const long e = fe.at(i, j, k);
            Ex[e] = ex; Ey[e] = ey; Ez[e] = ez;
        }
        __syncthreads();  // the next plane overwrites the current-plane arrays and ring slot (i + 1) & 1
    }
}

// i-planes per workgroup of the tiled EMF (one extra plane is read per run)
constexpr int EMF_IPLANES = 32;
static void emf3_tiled(const FSoa& FX, const FSoa& FY, const FSoa& FZ, const ERef& ref, const Ext& ce, const Ext& fe, const Box& eb,
                       int upwind, ET* Ex, ET* Ey, ET* Ez, hipStream_t s) {
    using T = EmfTile;
    const long nkt = (eb.hi[2] - eb.lo[2] + T::TK - 1) / T::TK, njt = ((eb.hi[1] - eb.lo[1] + T::TJ - 1) / T::TJ + 3) / 4 * 4;  // runs of 4
    const long nic = (eb.hi[0] - eb.lo[0] + EMF_IPLANES - 1) / EMF_IPLANES;
    if (nkt <= 0 || njt <= 0 || nic <= 0) return;
    hipLaunchKernelGGL(k_emf3_tiled, dim3((unsigned)(nkt * njt * nic)), dim3(256), 0, s, FX, FY, FZ, ref, ce, fe, eb, upwind, EMF_IPLANES,
                       Ex, Ey, Ez);
}

// Faraday (B_fields.cpp:21-66) fused with the clone: Bout = Bin everywhere,
// minus dt*curl(E) on the faces of the reference's ranges (ghost width gg).
template <int D, class BO>
__device__ __forceinline__ void faraday_body(const ET* Ex, const ET* Ey, const ET* Ez, const double* Bix, const double* Biy,
                                             const double* Biz, BO* Box_, BO* Boy, BO* Boz, const Ext& fe, int gg, double dt_dx,
                                             double dt_dy, double dt_dz, long m) {
    if (m >= fe.size) return;
    double bx = ld32(Bix, m), by = ld32(Biy, m), bz = ld32(Biz, m);
    // recover (i,j,k) of m
    unsigned r = (unsigned)m;
    int i, j = 0, k = 0;
    const unsigned Ny = D >= 2 ? fe.n[1] + 2 * fe.g : 1, Nz = D >= 3 ? fe.n[2] + 2 * fe.g : 1;
    if (D == 3) { k = (int)(r % Nz) - fe.g; r /= Nz; }
    if (D >= 2) { j = (int)(r % Ny) - fe.g; r /= Ny; }
    i = (int)r - fe.g;
    const int nx = fe.n[0] - 1, ny = fe.n[1] - 1, nz = fe.n[2] - 1;  // B sizes are n+1
    const bool ix = i >= -gg && i <= nx + gg, ixl = i >= -gg && i < nx + gg;
    const bool jy = j >= -gg && j <= ny + gg, jyl = j >= -gg && j < ny + gg;
    const bool kz = k >= -gg && k <= nz + gg, kzl = k >= -gg && k < nz + gg;
    if (D == 2) {
        if (ix && jyl) bx -= (ed(Ez, m + fe.s[1]) - ed(Ez, m)) * dt_dy;
        if (ixl && jy) by -= -(ed(Ez, m + fe.s[0]) - ed(Ez, m)) * dt_dx;
        if (ixl && jyl) bz -= (ed(Ey, m + fe.s[0]) - ed(Ey, m)) * dt_dx - (ed(Ex, m + fe.s[1]) - ed(Ex, m)) * dt_dy;
    } else {
        if (ix && jyl && kzl) bx -= (ed(Ez, m + fe.s[1]) - ed(Ez, m)) * dt_dy - (ed(Ey, m + fe.s[2]) - ed(Ey, m)) * dt_dz;
        if (ixl && jy && kzl) by -= (ed(Ex, m + fe.s[2]) - ed(Ex, m)) * dt_dz - (ed(Ez, m + fe.s[0]) - ed(Ez, m)) * dt_dx;
        if (ixl && jyl && kz) bz -= (ed(Ey, m + fe.s[0]) - ed(Ey, m)) * dt_dx - (ed(Ex, m + fe.s[1]) - ed(Ex, m)) * dt_dy;
    }
    Box_[m] = BO(bx); Boy[m] = BO(by); Boz[m] = BO(bz);
}
// The final update writes FP64 face B; the half-step one (gg = 1) writes the
// half-step faces in ET.
template <int D, class BO>
KB __global__ void k_faraday(const ET* Ex, const ET* Ey, const ET* Ez, const double* Bix, const double* Biy,
                             const double* Biz, BO* Box_, BO* Boy, BO* Boz, Ext fe, int gg, double dt_dx, double dt_dy,
                             double dt_dz) {
    faraday_body<D, BO>(Ex, Ey, Ez, Bix, Biy, Biz, Box_, Boy, Boz, fe, gg, dt_dx, dt_dy, dt_dz, tid_xcd());
}

// computeBodyField + shouldProtectThermal (B_fields.cpp:70-114, production
// beta-gated form), applied to a freshly updated conservative state.
__device__ __forceinline__ P<double> body_field(P<double> w, double cbx, double cby, double cbz, const Phys& ph) {
    bool protect;
    if (ph.ct_energy == 0) protect = false;
    else if (ph.ct_energy == 1) protect = true;
    else {
        double B2 = (w.Bx * w.Bx + w.By * w.By + w.Bz * w.Bz);
        protect = B2 == 0 ? false : (ph.eightpi * w.p / (w.Bx * w.Bx + w.By * w.By + w.Bz * w.Bz) <= ph.ct_beta);
    }
    if (!protect) {
        U<double> u = cons<true>(w, ph);
        u.Bx = cbx; u.By = cby; u.Bz = cbz;
        return prim<true>(u, ph);
    }
    w.Bx = cbx; w.By = cby; w.Bz = cbz;
    return w;
}
template <class T>
__device__ __forceinline__ void cell_B(const T* Bx, const T* By, const T* Bz, const Ext& fe, int D, int i, int j, int k,
                                       double& x, double& y, double& z) {
    const long m = fe.at(i, j, k);
    x = (ed(Bx, m) + ed(Bx, m + fe.s[0])) * 0.5;
    y = (ed(By, m) + ed(By, m + fe.s[1])) * 0.5;
    z = D == 3 ? (ed(Bz, m) + ed(Bz, m + fe.s[2])) * 0.5 : ed(Bz, m);
}

// Half-step state for the CTU corrector (CTU.cpp:57-61, 154-158): applyFluxes
// at dt/2 over cells [-1, n+1), body fields from the half-step faces, then
// Ehalf = cross(B, v) over cells [-1, n].
template <int D>
__device__ __forceinline__ void whalf_body(const Soa& W, const FSoa& FX, const FSoa& FY, const FSoa& FZ, const ET* Bhx, const ET* Bhy,
                                           const ET* Bhz, const Ext& ce, const Ext& fe, const Box& box, double hx, double hy, double hz,
                                           const Phys& ph, ET* Ehx, ET* Ehy, ET* Ehz, unsigned long long* err, long t) {
    int i, j, k;
    if (!box.point(t, i, j, k)) return;
    const long m = ce.at(i, j, k);
    P<double> w = load_p<double>(W, m, true);
    U<double> u = cons<true>(w, ph);
    u = uadd<true>(u, umul<true>(usub<true>(load_u_nb<0>(FX, m), load_u_nb<0>(FX, m + ce.s[0])), hx));
    u = uadd<true>(u, umul<true>(usub<true>(load_u_nb<1>(FY, m), load_u_nb<1>(FY, m + ce.s[1])), hy));
    if (D == 3) u = uadd<true>(u, umul<true>(usub<true>(load_u_nb<2>(FZ, m), load_u_nb<2>(FZ, m + ce.s[2])), hz));
    if (!finite<true>(u)) report(err, ST_WHALF_NAN, (unsigned long long)m, E_NAN);
    w = prim<true>(u, ph);
    double cbx, cby, cbz;
    cell_B(Bhx, Bhy, Bhz, fe, D, i, j, k, cbx, cby, cbz);
    w = body_field(w, cbx, cby, cbz, ph);
    double ex, ey, ez;
    cross(w.Bx, w.By, w.Bz, w.vx, w.vy, w.vz, ex, ey, ez);
    const long e = fe.at(i, j, k);
    Ehx[e] = ex; Ehy[e] = ey; Ehz[e] = ez;
}
template <int D>
KB __global__ void k_whalf(Soa W, FSoa FX, FSoa FY, FSoa FZ, const ET* Bhx, const ET* Bhy, const ET* Bhz, Ext ce,
                           Ext fe, Box box, double hx, double hy, double hz, Phys ph, ET* Ehx, ET* Ehy, ET* Ehz,
                           unsigned long long* err) {
    whalf_body<D>(W, FX, FY, FZ, Bhx, Bhy, Bhz, ce, fe, box, hx, hy, hz, ph, Ehx, Ehy, Ehz, err, tid_xcd());
}

// ctu_sweep_MHD corrections (CTU.cpp:65-129 3D, 162-202 2D), in place.
template <int D>
KB __global__ void k_ctu_mhd(Soa W, const double* Bx, const double* By, const double* Bz, const ET* Bhx, const ET* Bhy,
                             const ET* Bhz, const ET* E0x, const ET* E0y, const ET* E0z, FSoa FX, FSoa FY, FSoa FZ,
                             Soa XL, Soa XR, Soa YL, Soa YR, Soa ZL, Soa ZR, Ext ce, Ext fe, Box box, double dt_dx, double dt_dy,
                             double dt_dz, Phys ph) {
    int i, j, k;
    if (!box.point(tid(), i, j, k)) return;
    const long m = ce.at(i, j, k), q = fe.at(i, j, k);
    const long fx = fe.s[0], fy = fe.s[1], fz = fe.s[2];
    const double k8 = ph.k1_8pi;
    const P<double> Wc = load_p<double>(W, m, true);
    const double dBx = (Bx[q + fx] - Bx[q]) * dt_dx;
    const double dBy = (By[q + fy] - By[q]) * dt_dy;
    auto apply = [&](Soa& HL, Soa& HR, U<double> corr, int ax) {
        U<double> uL = uadd<true>(cons<true>(load_p<double>(HL, m, true), ph), corr);
        U<double> uR = uadd<true>(cons<true>(load_p<double>(HR, m, true), ph), corr);
        if (ax == 0) { uL.Bx = ed(Bhx, q); uR.Bx = ed(Bhx, q + fx); }
        else if (ax == 1) { uL.By = ed(Bhy, q); uR.By = ed(Bhy, q + fy); }
        else { uL.Bz = ed(Bhz, q); uR.Bz = ed(Bhz, q + fz); }
        store_p<double>(HL, m, prim<true>(uL, ph), true);
        store_p<double>(HR, m, prim<true>(uR, ph), true);
    };
    auto fd = [&](const FSoa& F, int ax) { return usub<true>(load_u(F, m, true), load_u(F, m + ce.s[ax], true)); };
    if (D == 2) {
        U<double> corr = umul<true>(umul<true>(fd(FY, 1), 0.5), dt_dy);
        corr.Bx = corr.By = corr.Bz = 0;
        double lim_z = minmod(-dBy, dBx);
        corr.E += k8 * (Wc.Bz * Wc.vz * lim_z);
        corr.Bz = 0.5 * dt_dy * (ed(E0x, q + fy) - ed(E0x, q)) + 0.5 * Wc.vz * lim_z;
        apply(XL, XR, corr, 0);
        corr = umul<true>(umul<true>(fd(FX, 0), 0.5), dt_dx);
        corr.Bx = corr.By = corr.Bz = 0;
        lim_z = minmod(-dBx, dBy);
        corr.E += k8 * (Wc.Bz * Wc.vz * lim_z);
        corr.Bz = -0.5 * dt_dx * (ed(E0y, q + fx) - ed(E0y, q)) + 0.5 * Wc.vz * lim_z;
        apply(YL, YR, corr, 1);
    } else {
        const double dBz = (Bz[q + fz] - Bz[q]) * dt_dz;
        // x half states
        U<double> corr = umul<true>(uadd<true>(umul<true>(fd(FY, 1), dt_dy), umul<true>(fd(FZ, 2), dt_dz)), 0.5);
        corr.Bx = corr.By = corr.Bz = 0;
        double lim_y = minmod(-dBz, dBx), lim_z = minmod(-dBy, dBx);
        corr.E += k8 * (Wc.By * Wc.vy * lim_y + Wc.Bz * Wc.vz * lim_z);
        corr.By = -0.25 * dt_dz * (ed(E0x, q + fz) - ed(E0x, q) + ed(E0x, q + fy + fz) - ed(E0x, q + fy)) + 0.5 * Wc.vy * lim_y;
        corr.Bz = 0.25 * dt_dy * (ed(E0x, q + fy) - ed(E0x, q) + ed(E0x, q + fy + fz) - ed(E0x, q + fz)) + 0.5 * Wc.vz * lim_z;
        apply(XL, XR, corr, 0);
        // y half states
        corr = umul<true>(uadd<true>(umul<true>(fd(FX, 0), dt_dx), umul<true>(fd(FZ, 2), dt_dz)), 0.5);
        corr.Bx = corr.By = corr.Bz = 0;
        double lim_x = minmod(-dBz, dBy);
        lim_z = minmod(-dBx, dBy);
        corr.E += k8 * (Wc.Bx * Wc.vx * lim_x + Wc.Bz * Wc.vz * lim_z);
        corr.Bx = 0.25 * dt_dz * (ed(E0y, q + fz) - ed(E0y, q) + ed(E0y, q + fx + fz) - ed(E0y, q + fx)) + 0.5 * Wc.vx * lim_x;
        corr.Bz = -0.25 * dt_dx * (ed(E0y, q + fx) - ed(E0y, q) + ed(E0y, q + fx + fz) - ed(E0y, q + fz)) + 0.5 * Wc.vz * lim_z;
        apply(YL, YR, corr, 1);
        // z half states
        corr = umul<true>(uadd<true>(umul<true>(fd(FX, 0), dt_dx), umul<true>(fd(FY, 1), dt_dy)), 0.5);
        corr.Bx = corr.By = corr.Bz = 0;
        lim_x = minmod(-dBy, dBz);
        lim_y = minmod(-dBx, dBz);
        corr.E += k8 * (Wc.Bx * Wc.vx * lim_x + Wc.By * Wc.vy * lim_y);
        corr.Bx = -0.25 * dt_dy * (ed(E0z, q + fy) - ed(E0z, q) + ed(E0z, q + fx + fy) - ed(E0z, q + fx)) + 0.5 * Wc.vx * lim_x;
        corr.By = 0.25 * dt_dx * (ed(E0z, q + fx) - ed(E0z, q) + ed(E0z, q + fx + fy) - ed(E0z, q + fy)) + 0.5 * Wc.vy * lim_y;
        apply(ZL, ZR, corr, 2);
    }
}

// ========================================================= apply + commit
// applyFluxes (Unsplit.cpp:256-288) + computeBodyFields with the new faces +
// the physicality check (Unsplit.cpp:94-99, 188-195), writing the complete
// next state: interior cells updated, ghost cells copied (the reference
// commits interior cells only, so the old ghosts remain).
// Returns the CFL signal speed of the new state for interior cells when
// `cfl` is set (0 elsewhere), for the fused CFL of the next step.
template <int D, bool MHD>
__device__ __forceinline__ double apply_body(const Soa& W, const Soa& Wn, const FSoa& FX, const FSoa& FY, const FSoa& FZ, const double* Bnx,
                                             const double* Bny, const double* Bnz, const Ext& ce, const Ext& fe, double dt_dx, double dt_dy,
                                             double dt_dz, const Phys& ph, unsigned long long* err, long m, bool cfl = false, int cfl_mode = 1,
                                             double cdx = 1, double cdy = 1, double cdz = 1) {
    if (m >= ce.size) return 0.0;
    const int nf = MHD ? 8 : 5;
    unsigned r = (unsigned)m;
    int i, j = 0, k = 0;
    const unsigned Ny = D >= 2 ? ce.n[1] + 2 * ce.g : 1, Nz = D >= 3 ? ce.n[2] + 2 * ce.g : 1;
    if (D == 3) { k = (int)(r % Nz) - ce.g; r /= Nz; }
    if (D >= 2) { j = (int)(r % Ny) - ce.g; r /= Ny; }
    i = (int)r - ce.g;
    const bool interior = i >= 0 && i < ce.n[0] && j >= 0 && j < ce.n[1] && k >= 0 && k < ce.n[2];
    if (!interior) {
        for (int f = 0; f < nf; f++) st_nt(Wn, f, m, W(f, m));
        return 0.0;
    }
    P<double> w = load_p<double>(W, m, MHD);
    U<double> u = cons<MHD>(w, ph);
    if constexpr (MHD) {
        u = uadd<true>(u, umul<true>(usub<true>(load_u_nb<0>(FX, m), load_u_nb<0>(FX, m + ce.s[0])), dt_dx));
        if (D >= 2) u = uadd<true>(u, umul<true>(usub<true>(load_u_nb<1>(FY, m), load_u_nb<1>(FY, m + ce.s[1])), dt_dy));
        if (D == 3) u = uadd<true>(u, umul<true>(usub<true>(load_u_nb<2>(FZ, m), load_u_nb<2>(FZ, m + ce.s[2])), dt_dz));
    } else {
        u = uadd<false>(u, umul<false>(usub<false>(load_u(FX, m, false), load_u(FX, m + ce.s[0], false)), dt_dx));
        if (D >= 2) u = uadd<false>(u, umul<false>(usub<false>(load_u(FY, m, false), load_u(FY, m + ce.s[1], false)), dt_dy));
        if (D == 3) u = uadd<false>(u, umul<false>(usub<false>(load_u(FZ, m, false), load_u(FZ, m + ce.s[2], false)), dt_dz));
    }
    if (D >= 2 && !finite<MHD>(u)) report(err, ST_APPLY_NAN, (unsigned long long)m, E_NAN);
    w = prim<MHD>(u, ph);
    if (MHD && D >= 2) {
        double cbx, cby, cbz;
        cell_B(Bnx, Bny, Bnz, fe, D, i, j, k, cbx, cby, cbz);
        w = body_field(w, cbx, cby, cbz, ph);
    }
    if (!physical<MHD>(w)) report(err, ST_UNPHYSICAL, (unsigned long long)m, E_UNPHYSICAL);
    store_p_nt(Wn, m, w, MHD);
    return cfl ? dmax(0.0, cfl_speed_w<D, MHD>(w, ph, cfl_mode, cdx, cdy, cdz)) : 0.0;  // NaN never wins
}
template <int D, bool MHD>
KB __global__ void k_apply(Soa W, Soa Wn, FSoa FX, FSoa FY, FSoa FZ, const double* Bnx, const double* Bny, const double* Bnz, Ext ce,
                           Ext fe, double dt_dx, double dt_dy, double dt_dz, Phys ph, unsigned long long* err, double* cflpart,
                           int cfl_mode, double cdx, double cdy, double cdz) {
    const double v = apply_body<D, MHD>(W, Wn, FX, FY, FZ, Bnx, Bny, Bnz, ce, fe, dt_dx, dt_dy, dt_dz, ph, err, tid_xcd(), cflpart != nullptr,
                                        cfl_mode, cdx, cdy, cdz);
    if (cflpart) block_max_store(v, cflpart);  // uniform branch
}

// 1D Godunov::sweep update (Splitting.cpp:67-95): cells [-g+1, n+g-2] take
// the flux difference; cell n+g-2 then takes the sweep's trailing zero
// update (a prim->cons->prim round trip).  Ghosts are committed too.
template <bool MHD>
KB __global__ void k_apply1(Soa W, Soa Wn, FSoa F, Ext ce, double dt_dx, Phys ph, unsigned long long* err) {
    const long m = tid();
    if (m >= ce.size) return;
    const int nf = MHD ? 8 : 5, g = ce.g, n = ce.n[0];
    const int i = (int)m - g;
    if (i < -g + 1 || i > n + g - 2) {
        for (int f = 0; f < nf; f++) Wn(f, m) = W(f, m);
        return;
    }
    P<double> w = load_p<double>(W, m, MHD);
    U<double> u = uadd<MHD>(cons<MHD>(w, ph), umul<MHD>(usub<MHD>(load_u(F, m, MHD), load_u(F, m + 1, MHD)), dt_dx));
    w = prim<MHD>(u, ph);
    if (i == n + g - 2) {
        U<double> fr = load_u(F, m + 1, MHD);
        u = uadd<MHD>(cons<MHD>(w, ph), umul<MHD>(usub<MHD>(fr, fr), dt_dx));
        w = prim<MHD>(u, ph);
    }
    if (i >= 0 && i < n && !physical<MHD>(w)) report(err, ST_UNPHYSICAL, (unsigned long long)m, E_UNPHYSICAL);
    store_p<double>(Wn, m, w, MHD);
}

// AoS <-> SoA staging for upload/download in the reference's record layout.
KB __global__ void k_aos_to_soa(const double* a, int stride, int nf, Soa s, long n) {
    const long m = tid();
    if (m >= n) return;
    for (int f = 0; f < nf; f++) s(f, m) = f < stride ? a[m * stride + f] : 0.0;
}
KB __global__ void k_soa_to_aos(Soa s, int nf, double* a, int stride, long n) {
    const long m = tid();
    if (m >= n) return;
    for (int f = 0; f < stride; f++) a[m * stride + f] = f < nf ? s(f, m) : 0.0;
}
KB __global__ void k_fill_u64(unsigned long long* p, unsigned long long v, int n) {
    const int t = (int)tid();
    if (t < n) p[t] = v;
}

// ============================================================== launchers
static inline unsigned blocks(long n) { return (unsigned)((n + 255) / 256); }

// initialize_B_fields (B_fields.cpp:117-149): cell B = face average on
// every cell, pressure held.
template <int D>
KB __global__ void k_init_B(Soa W, const double* Bx, const double* By, const double* Bz, Ext ce, Ext fe) {
    Box box = make_box(D, -ce.g, ce.n[0] + ce.g, -ce.g, ce.n[1] + ce.g, -ce.g, ce.n[2] + ce.g);
    int i, j, k;
    if (!box.point(tid(), i, j, k)) return;
    double x, y, z;
    cell_B(Bx, By, Bz, fe, D, i, j, k, x, y, z);
    const long m = ce.at(i, j, k);
    W(F_BX, m) = x; W(F_BY, m) = y; W(F_BZ, m) = z;
}
// Divergence monitor: max |div B| and max |B_cell| over interior cells.
template <int D>
KB __global__ void k_divb(const double* Bx, const double* By, const double* Bz, Ext ce, Ext fe, double dx, double dy, double dz,
                          unsigned long long* out) {
    Box box = make_box(D, 0, ce.n[0], 0, ce.n[1], 0, ce.n[2]);
    int i, j, k;
    double dv = 0, bm = 0;
    if (box.point(tid(), i, j, k)) {
        const long q = fe.at(i, j, k);
        double d = (Bx[q + fe.s[0]] - Bx[q]) / dx + (By[q + fe.s[1]] - By[q]) / dy;
        if (D == 3) d += (Bz[q + fe.s[2]] - Bz[q]) / dz;
        dv = dabs(d);
        double x, y, z;
        cell_B(Bx, By, Bz, fe, D, i, j, k, x, y, z);
        bm = dsqrt(x * x + y * y + z * z);
    }
    for (int off = 32; off > 0; off >>= 1) {
        dv = dmax(dv, __shfl_xor(dv, off));
        bm = dmax(bm, __shfl_xor(bm, off));
    }
    if ((threadIdx.x & 63) == 0) {
        if (dv > 0) atomicMax(out, (unsigned long long)__double_as_longlong(dv));
        if (bm > 0) atomicMax(out + 1, (unsigned long long)__double_as_longlong(bm));
    }
}
void launch_init_B(int dim, Soa W, double* B, Ext ce, Ext fe, hipStream_t s) {
    const long nf = fe.size;
    if (dim == 2) hipLaunchKernelGGL(k_init_B<2>, dim3(blocks(ce.size)), dim3(256), 0, s, W, B, B + nf, B + 2 * nf, ce, fe);
    else hipLaunchKernelGGL(k_init_B<3>, dim3(blocks(ce.size)), dim3(256), 0, s, W, B, B + nf, B + 2 * nf, ce, fe);
}
void launch_divb(int dim, double* B, Ext ce, Ext fe, const double d[3], unsigned long long* out, hipStream_t s) {
    const long nf = fe.size;
    Box b = make_box(dim, 0, ce.n[0], 0, ce.n[1], 0, ce.n[2]);
    if (dim == 2) hipLaunchKernelGGL(k_divb<2>, dim3(blocks(b.count())), dim3(256), 0, s, B, B + nf, B + 2 * nf, ce, fe, d[0], d[1], d[2], out);
    else hipLaunchKernelGGL(k_divb<3>, dim3(blocks(b.count())), dim3(256), 0, s, B, B + nf, B + 2 * nf, ce, fe, d[0], d[1], d[2], out);
}

void launch_bc(const BCArgs& a, bool faces, hipStream_t s) {
    long n = (long)(a.hi[0] - a.lo[0]) * (a.hi[1] - a.lo[1]);
    if (n <= 0) return;
    const bool contiguous = a.axis == a.dim - 1 && a.dim >= 2 && a.ce.n[a.axis] >= a.ce.g;
    if (faces && contiguous && a.kind == 1)
        hipLaunchKernelGGL(k_bc_faces_periodic_g, dim3(blocks(n * 2 * a.fe.g)), dim3(256), 0, s, a);
    else if (faces) hipLaunchKernelGGL(k_bc_faces, dim3(blocks(n)), dim3(256), 0, s, a);
    else if (contiguous)
        hipLaunchKernelGGL(k_bc_cells_g, dim3(blocks(n * (a.kind == 1 ? 2 : 1) * a.ce.g)), dim3(256), 0, s, a);
    else hipLaunchKernelGGL(k_bc_cells, dim3(blocks(n)), dim3(256), 0, s, a);
}

void launch_fill_u64(unsigned long long* p, unsigned long long v, int n, hipStream_t s) {
    hipLaunchKernelGGL(k_fill_u64, dim3(blocks(n)), dim3(256), 0, s, p, v, n);
}
void launch_aos_to_soa(const double* a, int stride, int nf, Soa soa, long n, hipStream_t s) {
    hipLaunchKernelGGL(k_aos_to_soa, dim3(blocks(n)), dim3(256), 0, s, a, stride, nf, soa, n);
}
void launch_soa_to_aos(Soa soa, int nf, double* a, int stride, long n, hipStream_t s) {
    hipLaunchKernelGGL(k_soa_to_aos, dim3(blocks(n)), dim3(256), 0, s, soa, nf, a, stride, n);
}

template <int D, bool MHD>
static void cfl_t(const StepCtx& c, unsigned long long* out, hipStream_t s) {
    const int n1 = c.ce.g < 1 ? c.ce.g : 1;
    Box b = make_box(D, -n1, c.ce.n[0] + n1, -n1, c.ce.n[1] + n1, -n1, c.ce.n[2] + n1);
    const unsigned nb = std::min(blocks(b.count()), 8192u);
    hipLaunchKernelGGL((k_cfl<D, MHD>), dim3(nb), dim3(256), 0, s, c.W, c.ce, b, c.ph, c.cfl_mode, c.d[0], c.d[1], c.d[2], out);
}
template <int D, bool MHD>
static void cfl_tail_t(const StepCtx& c, long nparts, unsigned long long* out, hipStream_t s) {
    const long work = nparts + (long)c.ce.n[0] * c.ce.n[1] * c.ce.n[2];  // upper bound; the kernel is grid-stride
    const unsigned nb = std::min(blocks(work), 2048u);
    hipLaunchKernelGGL((k_cfl_tail<D, MHD>), dim3(nb), dim3(256), 0, s, c.W, c.ce, c.cflpart, nparts, c.ph, c.cfl_mode, c.d[0], c.d[1],
                       c.d[2], out);
}
void launch_cfl_tail(const StepCtx& c, long nparts, unsigned long long* out, hipStream_t s) {
    if (c.dim == 2) c.mhd ? cfl_tail_t<2, true>(c, nparts, out, s) : cfl_tail_t<2, false>(c, nparts, out, s);
    else c.mhd ? cfl_tail_t<3, true>(c, nparts, out, s) : cfl_tail_t<3, false>(c, nparts, out, s);
}
void launch_cfl(const StepCtx& c, unsigned long long* out, hipStream_t s) {
    if (c.dim == 1) c.mhd ? cfl_t<1, true>(c, out, s) : cfl_t<1, false>(c, out, s);
    else if (c.dim == 2) c.mhd ? cfl_t<2, true>(c, out, s) : cfl_t<2, false>(c, out, s);
    else c.mhd ? cfl_t<3, true>(c, out, s) : cfl_t<3, false>(c, out, s);
}

template <int D, bool MHD, class R, int AX>
static void flux_t(const StepCtx& c, const Soa& HL, const Soa& HR, const FSoa& F, Box b, double dt_d, int stage, hipStream_t s) {
    if (b.count() <= 0) return;
    hipLaunchKernelGGL((k_flux<D, MHD, R, AX>), dim3(blocks(b.count())), dim3(256), 0, s, HL, HR, F, c.ce, b, dt_d, c.ph, c.err, stage);
}
template <int D, bool MHD, class R>
static void flux_ax(const StepCtx& c, int ax, const Soa& HL, const Soa& HR, const FSoa& F, Box b, double dt_d, int stage, hipStream_t s) {
    if (ax == 0) flux_t<D, MHD, R, 0>(c, HL, HR, F, b, dt_d, stage, s);
    else if (ax == 1) { if constexpr (D >= 2) flux_t<D, MHD, R, 1>(c, HL, HR, F, b, dt_d, stage, s); }
    else { if constexpr (D >= 3) flux_t<D, MHD, R, 2>(c, HL, HR, F, b, dt_d, stage, s); }
}

// Face-range boxes of computeFlux_* for axis `ax`: [lo, hi] inclusive along
// ax, [lo, hi) elsewhere; `w` is the transverse widening.
static Box face_box(int D, const Ext& ce, int ax, int along_lo, int along_hi_incl, int tlo, int thi_extra) {
    int lo[3], hi[3];
    for (int d = 0; d < 3; d++) {
        if (d == ax) { lo[d] = along_lo; hi[d] = ce.n[d] + along_hi_incl + 1; }
        else { lo[d] = tlo; hi[d] = ce.n[d] + thi_extra; }
    }
    return make_box(D, lo[0], hi[0], lo[1], hi[1], lo[2], hi[2]);
}

// One sweep launch for axis AX over the face box b.
template <int D, bool MHD, class R, bool FINAL, int AX>
static void sweep_t(const StepCtx& c, const Box& b, const FSoa& out, int stage, hipStream_t s) {
    SweepArgs a{};
    a.W = c.W;
    for (int q = 0; q < 3; q++) {
        a.B[q] = c.B[q]; a.Bh[q] = et(c.Bh[q]); a.E0[q] = et(c.E0[q]); a.Fp[q] = fs(c.Fp[q]);
        a.lo[q] = b.lo[q]; a.hi[q] = b.hi[q];
        a.dt_d[q] = c.dt / c.d[q];
    }
    a.Fo = out; a.ce = c.ce; a.fe = c.fe; a.ph = c.ph; a.err = c.err; a.stage = stage;
    if constexpr (eref_in_sweep<D, MHD, FINAL, AX>())
        if (stage == ST_PRED_X + AX)  // the CTU predictor (not the flux sweep of a step without CTU)
            for (int q = 0; q < 3; q++) a.eref[q] = et(c.Eh[q]);  // read by the predictor EMF before whalf overwrites it
    const long len = b.hi[AX] - b.lo[AX];
    long ntrans = 1;
    for (int d = 0; d < D; d++)
        if (d != AX) ntrans *= b.hi[d] - b.lo[d];
    if (len <= 0 || ntrans <= 0) return;
    if constexpr (AX == D - 1) {
        const long cells = ntrans * (len + 1);
        const long waves = cells > 1 ? (cells - 1 + 62) / 63 : 1;
        a.chunk = 63;
        hipLaunchKernelGGL((k_sweep_wave<D, MHD, R, AX, FINAL>), dim3(blocks(waves * 64)), dim3(256), 0, s, a);
    } else {
        // Walkers of at least 16 faces (one extra cell rebuild per walker),
        // ~2^22 of them where the grid allows: with 2 resident waves per SIMD
        // that is ~32 rounds of waves, so the last, partial round costs
        // little.  (A walker per whole line, 4161 waves at 512^3, ran as two
        // full rounds and a third of 65 waves.)
        const long want = 1L << 22;
        long nch = (want + ntrans - 1) / ntrans;
        long chunk = c.march_chunk > 0 ? c.march_chunk : (len + nch - 1) / nch;
        if (c.march_chunk <= 0 && chunk < 16) chunk = 16;
        if (chunk > len) chunk = len;
        nch = (len + chunk - 1) / chunk;
        a.chunk = (int)chunk;
        if constexpr (!march_window<R>()) {
            if (c.march_window) {
                hipLaunchKernelGGL((k_sweep_march_window<D, MHD, R, AX, FINAL>), dim3(blocks(nch * ntrans)), dim3(256), 0, s, a);
                return;
            }
        }
        hipLaunchKernelGGL((k_sweep_march<D, MHD, R, AX, FINAL>), dim3(blocks(nch * ntrans)), dim3(256), 0, s, a);
    }
}
template <int D, bool MHD, class R, bool FINAL>
static void sweep(const StepCtx& c, int ax, const Box& b, const FSoa& out, int stage, hipStream_t s) {
    if (ax == 0) sweep_t<D, MHD, R, FINAL, 0>(c, b, out, stage, s);
    else if (ax == 1) { if constexpr (D >= 2) sweep_t<D, MHD, R, FINAL, 1>(c, b, out, stage, s); }
    else { if constexpr (D >= 3) sweep_t<D, MHD, R, FINAL, 2>(c, b, out, stage, s); }
}

// The step with fused sweeps (MHD 2D/3D, hydro 2D): predictor fluxes ->
// [MHD: corner EMF, half-step faces, half-step reference EMF] -> corrector
// fluxes (CTU correction inside) -> [MHD: EMF, Faraday] -> apply.
template <int D, bool MHD, class R>
static void step_fused(const StepCtx& c, hipStream_t s) {
    const Ext& ce = c.ce;
    const double dt_d[3] = {c.dt / c.d[0], c.dt / c.d[1], c.dt / c.d[2]};
    if (c.ctu) {
        for (int ax = 0; ax < D; ax++) {
            Box b = MHD ? face_box(D, ce, ax, -1, 1, -2, 2) : face_box(D, ce, ax, 0, 0, -1, 1);
            sweep<D, MHD, R, false>(c, ax, b, fs(c.Fp[ax]), ST_PRED_X + ax, s);
            mark(c, K_PRED_X + ax);
        }
        if constexpr (MHD) {
            ERef none{nullptr, nullptr, nullptr};
            Box eb = make_box(D, -1, ce.n[0] + 2, -1, ce.n[1] + 2, -1, ce.n[2] + 2);
            if constexpr (D == 2 && eref_in_sweep<D, MHD, false, D - 1>())  // the reference EMF from the y sweep
                hipLaunchKernelGGL((k_emf2<false>), dim3(blocks(eb.count())), dim3(256), 0, s, fs(c.Fp[0]), fs(c.Fp[1]), c.W,
                                   ERef{et(c.Eh[0]), et(c.Eh[1]), et(c.Eh[2])}, ce, c.fe, eb, 1, et(c.E0[0]), et(c.E0[1]), et(c.E0[2]));
            else if constexpr (D == 2)
                hipLaunchKernelGGL((k_emf2<true>), dim3(blocks(eb.count())), dim3(256), 0, s, fs(c.Fp[0]), fs(c.Fp[1]), c.W, none, ce, c.fe, eb, 1,
                                   et(c.E0[0]), et(c.E0[1]), et(c.E0[2]));
            else {  // the reference EMF from the z sweep
                static_assert(eref_in_sweep<D, MHD, false, D - 1>(), "the 3D predictor sweep stores the reference EMF");
                emf3_tiled(fs(c.Fp[0]), fs(c.Fp[1]), fs(c.Fp[2]), ERef{et(c.Eh[0]), et(c.Eh[1]), et(c.Eh[2])}, ce, c.fe, eb, 1, et(c.E0[0]),
                           et(c.E0[1]), et(c.E0[2]), s);
            }
            mark(c, K_EMF_PRED);
            hipLaunchKernelGGL((k_faraday<D, ET>), dim3(blocks(c.fe.size)), dim3(256), 0, s, et(c.E0[0]), et(c.E0[1]), et(c.E0[2]), c.B[0], c.B[1],
                               c.B[2], et(c.Bh[0]), et(c.Bh[1]), et(c.Bh[2]), c.fe, 1, 0.5 * dt_d[0], 0.5 * dt_d[1], 0.5 * dt_d[2]);
            mark(c, K_FARADAY_HALF);
            Box cb = make_box(D, -1, ce.n[0] + 1, -1, ce.n[1] + 1, -1, ce.n[2] + 1);
            hipLaunchKernelGGL((k_whalf<D>), dim3(blocks(cb.count())), dim3(256), 0, s, c.W, fs(c.Fp[0]), fs(c.Fp[1]), fs(c.Fp[2]), et(c.Bh[0]),
                               et(c.Bh[1]), et(c.Bh[2]), ce, c.fe, cb, 0.5 * dt_d[0], 0.5 * dt_d[1], 0.5 * dt_d[2], c.ph, et(c.Eh[0]), et(c.Eh[1]),
                               et(c.Eh[2]), c.err);
            mark(c, K_WHALF);
        }
    }
    for (int ax = 0; ax < D; ax++) {
        Box b = MHD ? face_box(D, ce, ax, 0, 0, -1, 1) : face_box(D, ce, ax, 0, 0, 0, 0);
        if (c.ctu) sweep<D, MHD, R, true>(c, ax, b, fs(c.F[ax]), ST_FLUX_X + ax, s);
        else sweep<D, MHD, R, false>(c, ax, b, fs(c.F[ax]), ST_FLUX_X + ax, s);
        mark(c, K_FLUX_X + ax);
    }
    if constexpr (MHD) {
        Box eb = make_box(D, 0, ce.n[0] + 1, 0, ce.n[1] + 1, 0, ce.n[2] + 1);
        ERef eh{et(c.Eh[0]), et(c.Eh[1]), et(c.Eh[2])};
        if constexpr (D == 2)
            hipLaunchKernelGGL((k_emf2<false>), dim3(blocks(eb.count())), dim3(256), 0, s, fs(c.F[0]), fs(c.F[1]), c.W, eh, ce, c.fe, eb, c.ctu,
                               et(c.E0[0]), et(c.E0[1]), et(c.E0[2]));
        else
            emf3_tiled(fs(c.F[0]), fs(c.F[1]), fs(c.F[2]), eh, ce, c.fe, eb, c.ctu, et(c.E0[0]), et(c.E0[1]), et(c.E0[2]), s);
        mark(c, K_EMF);
        hipLaunchKernelGGL((k_faraday<D, double>), dim3(blocks(c.fe.size)), dim3(256), 0, s, et(c.E0[0]), et(c.E0[1]), et(c.E0[2]), c.B[0], c.B[1], c.B[2],
                           c.Bn[0], c.Bn[1], c.Bn[2], c.fe, 0, dt_d[0], dt_d[1], dt_d[2]);
        mark(c, K_FARADAY);
    }
    hipLaunchKernelGGL((k_apply<D, MHD>), dim3(blocks(ce.size)), dim3(256), 0, s, c.W, c.Wn, fs(c.F[0]), fs(c.F[1]), fs(c.F[2]), c.Bn[0], c.Bn[1],
                       c.Bn[2], ce, c.fe, dt_d[0], dt_d[1], dt_d[2], c.ph, c.err, c.cflpart, c.cfl_mode, c.d[0], c.d[1], c.d[2]);
    mark(c, K_APPLY);
}

template <int D, bool MHD, class R>
static void step_t(const StepCtx& c, hipStream_t s) {
    if constexpr ((MHD && D >= 2) || (!MHD && D == 2)) {
        if (c.fused) {
            step_fused<D, MHD, R>(c, s);
            return;
        }
    }
    const Ext& ce = c.ce;
    const double dt_d[3] = {c.dt / c.d[0], c.dt / c.d[1], c.dt / c.d[2]};
    const int g = ce.g;
    // half states
    {
        Box all = make_box(D, -g, ce.n[0] + g, -g, ce.n[1] + g, -g, ce.n[2] + g);
        hipLaunchKernelGGL((k_half<D, MHD, R>), dim3(blocks(all.count())), dim3(256), 0, s, c.W, c.B[0], c.B[1], c.B[2], ce, c.fe,
                           c.HS[0], c.HS[1], c.HS[2], c.HS[3], c.HS[4], c.HS[5], dt_d[0], dt_d[1], dt_d[2], c.ph);
        mark(c, K_HALF);
    }
    if (D == 1) {
        Box fb = make_box(1, -g + 1, ce.n[0] + g, 0, 1, 0, 1);
        flux_ax<D, MHD, R>(c, 0, c.HS[0], c.HS[1], fs(c.F[0]), fb, dt_d[0], ST_FLUX_X, s);
        mark(c, K_FLUX_X);
        if constexpr (D == 1)
            hipLaunchKernelGGL((k_apply1<MHD>), dim3(blocks(ce.size)), dim3(256), 0, s, c.W, c.Wn, fs(c.F[0]), ce, dt_d[0], c.ph, c.err);
        mark(c, K_APPLY);
        return;
    }
    if (c.ctu) {
        if constexpr (MHD) {
            // ctu_sweep_MHD predictor fluxes on [-1, n+1] x [-2, n+2)
            for (int ax = 0; ax < D; ax++) {
                flux_ax<D, MHD, R>(c, ax, c.HS[2 * ax], c.HS[2 * ax + 1], fs(c.F[ax]), face_box(D, ce, ax, -1, 1, -2, 2), dt_d[ax],
                                   ST_PRED_X + ax, s);
                mark(c, K_PRED_X + ax);
            }
            ERef none{nullptr, nullptr, nullptr};
            Box eb = make_box(D, -1, ce.n[0] + 2, -1, ce.n[1] + 2, -1, ce.n[2] + 2);
            if constexpr (D == 2)
                hipLaunchKernelGGL((k_emf2<true>), dim3(blocks(eb.count())), dim3(256), 0, s, fs(c.F[0]), fs(c.F[1]), c.W, none, ce, c.fe, eb, 1,
                                   et(c.E0[0]), et(c.E0[1]), et(c.E0[2]));
            else
                hipLaunchKernelGGL((k_emf3<true>), dim3(blocks(eb.count())), dim3(256), 0, s, fs(c.F[0]), fs(c.F[1]), fs(c.F[2]), c.W, none, ce, c.fe,
                                   eb, 1, et(c.E0[0]), et(c.E0[1]), et(c.E0[2]));
            mark(c, K_EMF_PRED);
            hipLaunchKernelGGL((k_faraday<D, ET>), dim3(blocks(c.fe.size)), dim3(256), 0, s, et(c.E0[0]), et(c.E0[1]), et(c.E0[2]), c.B[0], c.B[1],
                               c.B[2], et(c.Bh[0]), et(c.Bh[1]), et(c.Bh[2]), c.fe, 1, 0.5 * dt_d[0], 0.5 * dt_d[1], 0.5 * dt_d[2]);
            mark(c, K_FARADAY_HALF);
            Box cb = make_box(D, -1, ce.n[0] + 1, -1, ce.n[1] + 1, -1, ce.n[2] + 1);
            hipLaunchKernelGGL((k_whalf<D>), dim3(blocks(cb.count())), dim3(256), 0, s, c.W, fs(c.F[0]), fs(c.F[1]), fs(c.F[2]), et(c.Bh[0]), et(c.Bh[1]),
                               et(c.Bh[2]), ce, c.fe, cb, 0.5 * dt_d[0], 0.5 * dt_d[1], 0.5 * dt_d[2], c.ph, et(c.Eh[0]), et(c.Eh[1]), et(c.Eh[2]),
                               c.err);
            mark(c, K_WHALF);
            hipLaunchKernelGGL((k_ctu_mhd<D>), dim3(blocks(cb.count())), dim3(256), 0, s, c.W, c.B[0], c.B[1], c.B[2], et(c.Bh[0]), et(c.Bh[1]),
                               et(c.Bh[2]), et(c.E0[0]), et(c.E0[1]), et(c.E0[2]), fs(c.F[0]), fs(c.F[1]), fs(c.F[2]), c.HS[0], c.HS[1], c.HS[2], c.HS[3],
                               c.HS[4], c.HS[5], ce, c.fe, cb, dt_d[0], dt_d[1], dt_d[2], c.ph);
            mark(c, K_CTU);
        } else if constexpr (D == 2) {
            // ctu_sweep_hydro 2D (CTU.cpp:206-218)
            flux_ax<D, MHD, R>(c, 0, c.HS[0], c.HS[1], fs(c.F[0]), face_box(D, ce, 0, 0, 0, -1, 1), dt_d[0], ST_PRED_X, s);
            mark(c, K_PRED_X);
            flux_ax<D, MHD, R>(c, 1, c.HS[2], c.HS[3], fs(c.F[1]), face_box(D, ce, 1, 0, 0, -1, 1), dt_d[1], ST_PRED_Y, s);
            mark(c, K_PRED_Y);
            Box bx = make_box(2, -1, ce.n[0] + 1, 0, ce.n[1]);
            hipLaunchKernelGGL((k_correct<D, R>), dim3(blocks(bx.count())), dim3(256), 0, s, c.HS[0], c.HS[1], fs(c.F[1]), ce, bx, 1,
                               0.5 * dt_d[1], c.ph);
            mark(c, K_CTU);
            Box by = make_box(2, 0, ce.n[0], -1, ce.n[1] + 1);
            hipLaunchKernelGGL((k_correct<D, R>), dim3(blocks(by.count())), dim3(256), 0, s, c.HS[2], c.HS[3], fs(c.F[0]), ce, by, 0,
                               0.5 * dt_d[0], c.ph);
            mark(c, K_CTU);
        } else {
            ctu_hydro3(c, s);
            mark(c, K_CTU);
        }
    }
    // final fluxes
    if (MHD) {
        for (int ax = 0; ax < D; ax++) {
            flux_ax<D, MHD, R>(c, ax, c.HS[2 * ax], c.HS[2 * ax + 1], fs(c.F[ax]), face_box(D, ce, ax, 0, 0, -1, 1), dt_d[ax],
                               ST_FLUX_X + ax, s);
            mark(c, K_FLUX_X + ax);
        }
    } else {
        for (int ax = 0; ax < D; ax++) {
            flux_ax<D, MHD, R>(c, ax, c.HS[2 * ax], c.HS[2 * ax + 1], fs(c.F[ax]), face_box(D, ce, ax, 0, 0, 0, 0), dt_d[ax],
                               ST_FLUX_X + ax, s);
            mark(c, K_FLUX_X + ax);
        }
    }
    if constexpr (MHD) {
        Box eb = make_box(D, 0, ce.n[0] + 1, 0, ce.n[1] + 1, 0, ce.n[2] + 1);
        ERef eh{et(c.Eh[0]), et(c.Eh[1]), et(c.Eh[2])};
        if constexpr (D == 2)
            hipLaunchKernelGGL((k_emf2<false>), dim3(blocks(eb.count())), dim3(256), 0, s, fs(c.F[0]), fs(c.F[1]), c.W, eh, ce, c.fe, eb, c.ctu,
                               et(c.E0[0]), et(c.E0[1]), et(c.E0[2]));
        else
            hipLaunchKernelGGL((k_emf3<false>), dim3(blocks(eb.count())), dim3(256), 0, s, fs(c.F[0]), fs(c.F[1]), fs(c.F[2]), c.W, eh, ce, c.fe, eb,
                               c.ctu, et(c.E0[0]), et(c.E0[1]), et(c.E0[2]));
        mark(c, K_EMF);
        hipLaunchKernelGGL((k_faraday<D, double>), dim3(blocks(c.fe.size)), dim3(256), 0, s, et(c.E0[0]), et(c.E0[1]), et(c.E0[2]), c.B[0], c.B[1], c.B[2],
                           c.Bn[0], c.Bn[1], c.Bn[2], c.fe, 0, dt_d[0], dt_d[1], dt_d[2]);
        mark(c, K_FARADAY);
    }
    hipLaunchKernelGGL((k_apply<D, MHD>), dim3(blocks(ce.size)), dim3(256), 0, s, c.W, c.Wn, fs(c.F[0]), fs(c.F[1]), fs(c.F[2]), c.Bn[0], c.Bn[1],
                       c.Bn[2], ce, c.fe, dt_d[0], dt_d[1], dt_d[2], c.ph, c.err, c.cflpart, c.cfl_mode, c.d[0], c.d[1], c.d[2]);
    mark(c, K_APPLY);
}

// 3D hydro CTU, 12 solves (CTU.cpp:222-261): six corner-flux passes with
// dt/3 transverse corrections, then the state corrections.
void ctu_hydro3(const StepCtx& c, hipStream_t s) {
    const Ext& ce = c.ce;
    const double dx = c.dt / c.d[0], dy = c.dt / c.d[1], dz = c.dt / c.d[2];
    const Soa &xL = c.HS[0], &xR = c.HS[1], &yL = c.HS[2], &yR = c.HS[3], &zL = c.HS[4], &zR = c.HS[5];
    const FSoa FX = fs(c.F[0]), FY = fs(c.F[1]), FZ = fs(c.F[2]);
    FSoa T[5];  // five scratch flux arrays
    for (int q = 0; q < 5; q++) T[q] = fs(c.T[q]);
    const int nx = ce.n[0], ny = ce.n[1], nz = ce.n[2];
    flux_t<3, false, double, 0>(c, xL, xR, FX, make_box(3, 0, nx + 1, -1, ny + 1, -1, nz + 1), dx, ST_PRED_X, s);
    flux_t<3, false, double, 1>(c, yL, yR, FY, make_box(3, -1, nx + 1, 0, ny + 1, -1, nz + 1), dy, ST_PRED_Y, s);
    flux_t<3, false, double, 2>(c, zL, zR, FZ, make_box(3, -1, nx + 1, -1, ny + 1, 0, nz + 1), dz, ST_PRED_Z, s);
    auto ctu = [&](int ax, const Soa& L, const Soa& Rr, const FSoa& Tr, int tax, const FSoa& out, double dd, double dtt, int stage) {
        int isX = tax == 0, isY = tax == 1, isZ = tax == 2;
        Box b;
        if (ax == 0) b = make_box(3, 0, nx + 1, -1 + isY, ny + 1 - isY, -1 + isZ, nz + 1 - isZ);
        else if (ax == 1) b = make_box(3, -1 + isX, nx + 1 - isX, 0, ny + 1, -1 + isZ, nz + 1 - isZ);
        else b = make_box(3, -1 + isX, nx + 1 - isX, -1 + isY, ny + 1 - isY, 0, nz + 1);
        if (ax == 0)
            hipLaunchKernelGGL((k_ctu_flux<double, 0>), dim3(blocks(b.count())), dim3(256), 0, s, L, Rr, Tr, out, ce, b, tax, dd, dtt, c.ph, c.err, stage);
        else if (ax == 1)
            hipLaunchKernelGGL((k_ctu_flux<double, 1>), dim3(blocks(b.count())), dim3(256), 0, s, L, Rr, Tr, out, ce, b, tax, dd, dtt, c.ph, c.err, stage);
        else
            hipLaunchKernelGGL((k_ctu_flux<double, 2>), dim3(blocks(b.count())), dim3(256), 0, s, L, Rr, Tr, out, ce, b, tax, dd, dtt, c.ph, c.err, stage);
    };
    auto corr = [&](const Soa& L, const Soa& Rr, const FSoa& F, double dd, int sdim, int fdim) {
        int lo[3] = {0, 0, 0}, hi[3] = {nx, ny, nz};
        lo[sdim] = -1; hi[sdim] += 1;
        Box b = make_box(3, lo[0], hi[0], lo[1], hi[1], lo[2], hi[2]);
        hipLaunchKernelGGL((k_correct<3, double>), dim3(blocks(b.count())), dim3(256), 0, s, L, Rr, F, ce, b, fdim, dd, c.ph);
    };
    const FSoa &FXz = T[0], &FYz = T[1], &FXy = T[2], &FZy = T[3], &FYx = T[4];
    ctu(0, xL, xR, FZ, 2, FXz, dx, dz / 3.0, ST_CTU_F0 + 0);
    ctu(1, yL, yR, FZ, 2, FYz, dy, dz / 3.0, ST_CTU_F0 + 1);
    ctu(0, xL, xR, FY, 1, FXy, dx, dy / 3.0, ST_CTU_F0 + 2);
    ctu(2, zL, zR, FY, 1, FZy, dz, dy / 3.0, ST_CTU_F0 + 3);
    corr(xL, xR, FYz, 0.5 * dy, 0, 1);
    corr(xL, xR, FZy, 0.5 * dz, 0, 2);
    ctu(1, yL, yR, FX, 0, FYx, dy, dx / 3.0, ST_CTU_F0 + 4);
    const FSoa& FZx = FZ;  // FZ is no longer needed
    ctu(2, zL, zR, FX, 0, FZx, dz, dx / 3.0, ST_CTU_F0 + 5);
    corr(yL, yR, FXz, 0.5 * dx, 1, 0);
    corr(yL, yR, FZx, 0.5 * dz, 1, 2);
    corr(zL, zR, FXy, 0.5 * dx, 2, 0);
    corr(zL, zR, FYx, 0.5 * dy, 2, 1);
}

// ===================================================== batched patches
// Many independent 2D MHD patches (CTU + CT) advanced stage by stage: each
// launch below covers every active patch through a block table, and runs the
// same kernel body as the single-grid step over the patch's own arrays.  The
// step is begin -> boundary passes -> CFL -> dt -> predictor sweeps -> EMF ->
// half-step faces -> half-step state -> corrector sweeps -> EMF -> Faraday ->
// apply -> commit, a fixed sequence a graph can replay.  A patch whose
// attempt fails keeps its state and retries at half the dt in the next step
// (the dt halving of Grid::advance, CFL.cpp:151-213, one attempt per step).
struct PatchView {
    Ext ce, fe;
    Soa W, Wn;
    double *B[3], *Bn[3];
    ET* Bh[3];  // half-step faces
    ET *E0[3], *Eh[3];
    FSoa Fp[3], F[3];
    double dx, dy;
    int cls;
};
__device__ __forceinline__ PatchView patch_view(const BatchArgs& A, int p) {
    const PatchDesc P = A.P[p];
    constexpr int G = 3;  // validGhosts for MHD + CTU
    PatchView v;
    v.ce = make_ext(2, P.n[0], P.n[1], 1, G);
    v.fe = make_ext(2, P.n[0] + 1, P.n[1] + 1, 1, G);
    const int par = A.parity[p];
    v.W = Soa{A.W[par] + 8 * P.oc, P.nc};
    v.Wn = Soa{A.W[1 - par] + 8 * P.oc, P.nc};
    for (int c = 0; c < 3; c++) {
        v.B[c] = A.B[par] + 3 * P.of + c * P.nfc;
        v.Bn[c] = A.B[1 - par] + 3 * P.of + c * P.nfc;
        v.Bh[c] = et(A.Bh) + 3 * P.of + c * P.nfc;
        v.E0[c] = et(A.E0) + 3 * P.of + c * P.nfc;  // EMF pools hold ET elements at the face offsets
        v.Eh[c] = et(A.Eh) + 3 * P.of + c * P.nfc;
    }
    for (int ax = 0; ax < 2; ax++) {  // flux pools hold FT elements at the same per-patch offsets
        v.Fp[ax] = FSoa{reinterpret_cast<FT*>(A.Fp) + 16 * P.oc + ax * 8 * P.nc, P.nc};
        v.F[ax] = FSoa{reinterpret_cast<FT*>(A.F) + 16 * P.oc + ax * 8 * P.nc, P.nc};
    }
    v.Fp[2] = v.Fp[0];  // no z axis in 2D; never read
    v.F[2] = v.F[0];
    v.dx = P.d[0];
    v.dy = P.d[1];
    v.cls = P.cls;
    return v;
}
// workgroup -> (patch, work item), in the XCD order of block_xcd; inactive
// patches' workgroups leave at once
#define BATCH_ITEM(bt)                          \
    const unsigned bx = block_xcd();            \
    const int p = bt.patch[bx];                 \
    if (!A.active[p]) return;                   \
    const long t = (long)bt.first[bx] + threadIdx.x;

KB __global__ void kb_begin(BatchArgs A) {
    const int p = (int)tid();
    if (p >= A.npatch) return;
    const bool on = A.status[p] == 0 && (A.cls_filter < 0 || A.P[p].cls == A.cls_filter);
    A.active[p] = on ? 1 : 0;
    if (on) {
        // the interior's CFL maximum from the apply pass that wrote the
        // current state, when it did; kb_cfl adds the ghost shell (or, without
        // it, every cell)
        A.vmax[p] = A.cfl_ok[p] ? A.vnext[p] : 0ull;
        A.vnext[p] = 0ull;
        A.err[p] = ~0ull;
    }
}
// Boundary pass of every patch: both sides of one axis in one launch (they
// touch disjoint ghost slabs for sides of 4 cells or more), in the pass
// order of a list element's apply.
template <bool FACES, int AX>
KB __global__ void kb_bc(BatchArgs A, BlockTable bt) {
    BATCH_ITEM(bt)
    const PatchView v = patch_view(A, p);
    constexpr int G = 3;
    const long cnt = v.ce.n[AX == 0 ? 1 : 0] + 2 * G + (FACES ? 1 : 0);
    const bool periodic = A.bc_kind == 1;
    if (t >= (!FACES && AX == 1 ? 2 * G : periodic ? 1 : 2) * cnt) return;
    BCArgs b{};
    b.W = v.W;
    for (int c = 0; c < 3; c++) b.B[c] = v.B[c];
    b.ce = v.ce; b.fe = v.fe; b.dim = 2; b.nf = 8; b.kind = A.bc_kind; b.corners = 1; b.gated = 0;
    b.axis = AX;
    b.lo[0] = -G; b.hi[0] = (int)(-G + cnt);
    b.lo[1] = 0; b.hi[1] = 1;
    if constexpr (!FACES && AX == 1) {
        // The cell pass along y (the contiguous axis), one item per ghost
        // cell, both sides: patches hold at least 4 cells along y, so every
        // cell a ghost reads is interior and the g loop has no order to keep.
        const int side = (int)(t / (cnt * G));
        const long r = t % (cnt * G);
        bc_cell(b, b.lo[0] + (int)(r / G), 0, (int)(r % G) + 1, side);
        return;
    }
    b.side = periodic ? 2 : (int)(t / cnt);
    if (FACES) bc_faces_body(b, t % cnt);
    else bc_cells_body(b, t % cnt);
}
KB __global__ void kb_init_B(BatchArgs A, BlockTable bt) {
    BATCH_ITEM(bt)
    const PatchView v = patch_view(A, p);
    if (t >= v.ce.size) return;
    const int ny = v.ce.n[1] + 6;
    const int i = (int)(t / ny) - 3, j = (int)(t % ny) - 3;
    double x, y, z;
    cell_B(v.B[0], v.B[1], v.B[2], v.fe, 2, i, j, 0, x, y, z);
    v.W(F_BX, t) = x; v.W(F_BY, t) = y; v.W(F_BZ, t) = z;
}
KB __global__ void kb_cfl(BatchArgs A, BlockTable bt) {
    BATCH_ITEM(bt)
    const PatchView v = patch_view(A, p);
    const long ny2 = v.ce.n[1] + 2;
    double m = 0;
    const int i = (int)(t / ny2) - 1, j = (int)(t % ny2) - 1;
    const bool interior = i >= 0 && i < v.ce.n[0] && j >= 0 && j < v.ce.n[1];
    if (t < (long)(v.ce.n[0] + 2) * ny2 && !(interior && A.cfl_ok[p]))
        m = dmax(m, cfl_speed_at<2, true>(v.W, v.ce, i, j, 0, A.ph, A.cfl_mode, v.dx, v.dy, 1.0));
    for (int off = 32; off > 0; off >>= 1) m = dmax(m, __shfl_xor(m, off));
    __shared__ double wmax[BATCH_THREADS / 64];
    if ((threadIdx.x & 63) == 0) wmax[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        double b = wmax[0];
        for (int q = 1; q < BATCH_THREADS / 64; q++) b = dmax(b, wmax[q]);
        if (b > 0) atomicMax(A.vmax + p, (unsigned long long)__double_as_longlong(b));
    }
}
KB __global__ void kb_dt(BatchArgs A) {
    const int p = (int)tid();
    if (p >= A.npatch || !A.active[p]) return;
    const double m = __longlong_as_double((long long)A.vmax[p]);
    if (!(m > 0)) {  // "CFL timestep failed: max signal speed is zero"
        A.active[p] = 0;
        if (A.commit) A.status[p] = BATCH_ERR_CFL;
        return;
    }
    double dt = A.cfl_coeff / m;
    for (int h = 0; h < A.halv[p]; h++) dt *= 0.5;
    A.dt[p] = dt;
}
template <int AX, bool FIN>
__device__ __forceinline__ SweepArgs batch_sweep_args(const BatchArgs& A, const PatchView& v, int p) {
    SweepArgs s{};
    const double dt = A.dt[p];
    const double dt_d[3] = {dt / v.dx, dt / v.dy, dt / 1.0};
    s.W = v.W;
    for (int q = 0; q < 3; q++) {
        s.B[q] = v.B[q]; s.Bh[q] = v.Bh[q]; s.E0[q] = v.E0[q]; s.Fp[q] = v.Fp[q]; s.dt_d[q] = dt_d[q];
    }
    s.ce = v.ce; s.fe = v.fe; s.ph = A.ph; s.err = A.err + p;
    const int nx = v.ce.n[0], ny = v.ce.n[1];
    constexpr int T = AX == 0 ? 1 : 0;
    // predictor faces [-1, n+1] x [-2, n+2); corrector faces [0, n] x [-1, n+1)
    s.lo[AX] = FIN ? 0 : -1;
    s.hi[AX] = (AX == 0 ? nx : ny) + (FIN ? 1 : 2);
    s.lo[T] = FIN ? -1 : -2;
    s.hi[T] = (T == 0 ? nx : ny) + (FIN ? 1 : 2);
    s.lo[2] = 0; s.hi[2] = 1;
    s.Fo = FIN ? v.F[AX] : v.Fp[AX];
    s.stage = (FIN ? ST_FLUX_X : ST_PRED_X) + AX;
    if constexpr (eref_in_sweep<2, true, FIN, AX>())
        for (int q = 0; q < 3; q++) s.eref[q] = v.Eh[q];  // the predictor EMF's reference, before whalf overwrites it
    return s;
}
template <bool FIN>
KS __global__ void kb_sweep_x(BatchArgs A, BlockTable bt) {
    BATCH_ITEM(bt)
    const PatchView v = patch_view(A, p);
    SweepArgs s = batch_sweep_args<0, FIN>(A, v, p);
    const int len = s.hi[0] - s.lo[0];
    s.chunk = A.chunk[v.cls] < len ? A.chunk[v.cls] : len;
    sweep_march_body<2, true, RT, 0, FIN, march_window<RT>(), true>(s, t);
}
template <bool FIN>
KS __attribute__((amdgpu_waves_per_eu(wave_min_waves<RT>()))) __global__ void kb_sweep_y(BatchArgs A, BlockTable bt) {
    BATCH_ITEM(bt)  // first is a multiple of the workgroup size: first / 64 + the wavefront in the workgroup
    const PatchView v = patch_view(A, p);
    SweepArgs s = batch_sweep_args<1, FIN>(A, v, p);
    const long w = (long)bt.first[bx] / 64 + __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int lane = threadIdx.x & 63;
    switch (A.ysw[v.cls]) {  // segment width of the patch's size class (uniform per workgroup)
        case 0: s.chunk = 63; sweep_stream_body<2, true, RT, 1, FIN>(s, (int)w, lane); break;  // the patch's columns streamed
        case 16: s.chunk = 15; sweep_wave_body<2, true, RT, 1, FIN, 16>(s, w * 4 + lane / 16, lane); break;
        case 32: s.chunk = 31; sweep_wave_body<2, true, RT, 1, FIN, 32>(s, w * 2 + lane / 32, lane); break;
        default: s.chunk = 63; sweep_wave_body<2, true, RT, 1, FIN, 64>(s, w, lane); break;
    }
}
template <bool PRED>
KB __global__ void kb_emf(BatchArgs A, BlockTable bt) {
    BATCH_ITEM(bt)
    const PatchView v = patch_view(A, p);
    const int lo = PRED ? -1 : 0, ex = PRED ? 2 : 1;
    const Box eb = make_box(2, lo, v.ce.n[0] + ex, lo, v.ce.n[1] + ex);
    if (PRED) {
        if constexpr (eref_in_sweep<2, true, false, 1>()) {  // the reference EMF from the y sweep
            const ERef eh{v.Eh[0], v.Eh[1], v.Eh[2]};
            emf2_body<false>(v.Fp[0], v.Fp[1], v.W, eh, v.ce, v.fe, eb, 1, v.E0[0], v.E0[1], v.E0[2], t);
        } else {
            const ERef none{nullptr, nullptr, nullptr};
            emf2_body<true>(v.Fp[0], v.Fp[1], v.W, none, v.ce, v.fe, eb, 1, v.E0[0], v.E0[1], v.E0[2], t);
        }
    } else {
        const ERef eh{v.Eh[0], v.Eh[1], v.Eh[2]};
        emf2_body<false>(v.F[0], v.F[1], v.W, eh, v.ce, v.fe, eb, 1, v.E0[0], v.E0[1], v.E0[2], t);
    }
}
template <bool HALF>
KB __global__ void kb_faraday(BatchArgs A, BlockTable bt) {
    BATCH_ITEM(bt)
    const PatchView v = patch_view(A, p);
    const double dt = A.dt[p], f = HALF ? 0.5 : 1.0;
    const double dx = f * (dt / v.dx), dy = f * (dt / v.dy), dz = f * (dt / 1.0);
    if (HALF) faraday_body<2, ET>(v.E0[0], v.E0[1], v.E0[2], v.B[0], v.B[1], v.B[2], v.Bh[0], v.Bh[1], v.Bh[2], v.fe, 1, dx, dy, dz, t);
    else faraday_body<2, double>(v.E0[0], v.E0[1], v.E0[2], v.B[0], v.B[1], v.B[2], v.Bn[0], v.Bn[1], v.Bn[2], v.fe, 0, dx, dy, dz, t);
}
KB __global__ void kb_whalf(BatchArgs A, BlockTable bt) {
    BATCH_ITEM(bt)
    const PatchView v = patch_view(A, p);
    const double dt = A.dt[p];
    const Box cb = make_box(2, -1, v.ce.n[0] + 1, -1, v.ce.n[1] + 1);
    whalf_body<2>(v.W, v.Fp[0], v.Fp[1], v.Fp[2], v.Bh[0], v.Bh[1], v.Bh[2], v.ce, v.fe, cb, 0.5 * (dt / v.dx), 0.5 * (dt / v.dy),
                  0.5 * (dt / 1.0), A.ph, v.Eh[0], v.Eh[1], v.Eh[2], A.err + p, t);
}
KB __global__ void kb_apply(BatchArgs A, BlockTable bt) {
    BATCH_ITEM(bt)
    const PatchView v = patch_view(A, p);
    const double dt = A.dt[p];
    double m = apply_body<2, true>(v.W, v.Wn, v.F[0], v.F[1], v.F[2], v.Bn[0], v.Bn[1], v.Bn[2], v.ce, v.fe, dt / v.dx, dt / v.dy,
                                   dt / 1.0, A.ph, A.err + p, t, true, A.cfl_mode, v.dx, v.dy, 1.0);
    // the new state's CFL maximum over the patch's interior, for the next
    // step if this one is committed (one atomic per workgroup)
    for (int off = 32; off > 0; off >>= 1) m = dmax(m, __shfl_xor(m, off));
    __shared__ double wmax[BATCH_THREADS / 64];
    if ((threadIdx.x & 63) == 0) wmax[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        double b = wmax[0];
        for (int q = 1; q < BATCH_THREADS / 64; q++) b = dmax(b, wmax[q]);
        if (b > 0) atomicMax(A.vnext + p, (unsigned long long)__double_as_longlong(b));
    }
}
KB __global__ void kb_commit(BatchArgs A) {
    const int p = (int)tid();
    if (p >= A.npatch || !A.active[p]) return;
    // vnext describes the state only when the attempt was committed
    A.cfl_ok[p] = A.commit && A.err[p] == ~0ull ? 1 : 0;
    if (!A.commit) return;
    if (A.err[p] == ~0ull) {
        A.parity[p] = 1 - A.parity[p];
        A.dt_last[p] = A.dt[p];
        A.steps[p] += 1;
        A.halv[p] = 0;
    } else {
        A.fails[p] += 1;
        A.halv[p] += 1;
        if (0.5 * A.dt[p] < A.tol) A.status[p] = BATCH_ERR_DT;  // "Timestep has fallen below minimum"
    }
}

void launch_batch_step(const BatchArgs& a, const BlockTable* bt, bool init_only, hipStream_t s) {
    const dim3 T(BATCH_THREADS);
    const unsigned pb = blocks(a.npatch);
    auto grid = [&](int k) { return dim3(bt[k].nblocks); };
    hipLaunchKernelGGL(kb_begin, dim3(pb), T, 0, s, a);
    hipLaunchKernelGGL((kb_bc<false, 0>), grid(BK_BCC_X), T, 0, s, a, bt[BK_BCC_X]);
    hipLaunchKernelGGL((kb_bc<false, 1>), grid(BK_BCC_Y), T, 0, s, a, bt[BK_BCC_Y]);
    hipLaunchKernelGGL((kb_bc<true, 0>), grid(BK_BCF_X), T, 0, s, a, bt[BK_BCF_X]);
    hipLaunchKernelGGL((kb_bc<true, 1>), grid(BK_BCF_Y), T, 0, s, a, bt[BK_BCF_Y]);
    if (init_only) {
        hipLaunchKernelGGL(kb_init_B, grid(BK_CELLS), T, 0, s, a, bt[BK_CELLS]);
        return;
    }
    hipLaunchKernelGGL(kb_cfl, grid(BK_CELLS1), T, 0, s, a, bt[BK_CELLS1]);
    hipLaunchKernelGGL(kb_dt, dim3(pb), T, 0, s, a);
    hipLaunchKernelGGL((kb_sweep_x<false>), grid(BK_PRED_X), T, 0, s, a, bt[BK_PRED_X]);
    hipLaunchKernelGGL((kb_sweep_y<false>), grid(BK_PRED_Y), T, 0, s, a, bt[BK_PRED_Y]);
    hipLaunchKernelGGL((kb_emf<true>), grid(BK_EDGES_P), T, 0, s, a, bt[BK_EDGES_P]);
    hipLaunchKernelGGL((kb_faraday<true>), grid(BK_FACES), T, 0, s, a, bt[BK_FACES]);
    hipLaunchKernelGGL(kb_whalf, grid(BK_CELLS1), T, 0, s, a, bt[BK_CELLS1]);
    hipLaunchKernelGGL((kb_sweep_x<true>), grid(BK_FLUX_X), T, 0, s, a, bt[BK_FLUX_X]);
    hipLaunchKernelGGL((kb_sweep_y<true>), grid(BK_FLUX_Y), T, 0, s, a, bt[BK_FLUX_Y]);
    hipLaunchKernelGGL((kb_emf<false>), grid(BK_EDGES), T, 0, s, a, bt[BK_EDGES]);
    hipLaunchKernelGGL((kb_faraday<false>), grid(BK_FACES), T, 0, s, a, bt[BK_FACES]);
    hipLaunchKernelGGL(kb_apply, grid(BK_CELLS), T, 0, s, a, bt[BK_CELLS]);
    hipLaunchKernelGGL(kb_commit, dim3(pb), T, 0, s, a);
}

// ====================================================== Strang split step
// One directional sweep of Grid2D/3D::split_step (advanceX/Y/Z,
// Splitting.cpp:219-299): Godunov::sweep (Splitting.cpp:67-95) along every
// line of axis AX, ghost lines included.  The faces come from the fused sweep
// body over faces [-g+1, n+g-1] (first order at both ends, as the 1D sweep);
// the update runs in the line's own frame (XY / XZ swap), where the
// reference's y and z sweeps do their prim -> cons -> prim arithmetic, and
// cell n+g-2 takes the sweep's trailing zero update.  Hydro only.
template <class T> __device__ __forceinline__ P<T> to_line(const P<T>& w, int ax) { return ax == 0 ? w : ax == 1 ? swapXY(w) : swapXZ(w); }
template <class T> __device__ __forceinline__ U<T> to_line(const U<T>& u, int ax) { return ax == 0 ? u : ax == 1 ? swapXY(u) : swapXZ(u); }
template <int D, int AX>
KB __global__ void k_apply_axis(Soa W, Soa Wn, FSoa F, Ext ce, double dt_d, Phys ph) {
    const long m = tid();
    if (m >= ce.size) return;
    const int g = ce.g, n = ce.n[AX];
    unsigned r = (unsigned)m;
    int idx[3] = {0, 0, 0};
    const unsigned Ny = D >= 2 ? ce.n[1] + 2 * g : 1, Nz = D >= 3 ? ce.n[2] + 2 * g : 1;
    if (D == 3) { idx[2] = (int)(r % Nz) - g; r /= Nz; }
    if (D >= 2) { idx[1] = (int)(r % Ny) - g; r /= Ny; }
    idx[0] = (int)r - g;
    const int a = idx[AX];
    if (a < -g + 1 || a > n + g - 2) {
        for (int f = 0; f < 5; f++) Wn(f, m) = W(f, m);
        return;
    }
    P<double> w = to_line(load_p<double>(W, m, false), AX);
    const U<double> fl = to_line(load_u(F, m, false), AX), fr = to_line(load_u(F, m + ce.s[AX], false), AX);
    U<double> u = uadd<false>(cons<false>(w, ph), umul<false>(usub<false>(fl, fr), dt_d));
    w = prim<false>(u, ph);
    if (a == n + g - 2) {  // the sweep's last update: the same flux on both sides
        u = uadd<false>(cons<false>(w, ph), umul<false>(usub<false>(fr, fr), dt_d));
        w = prim<false>(u, ph);
    }
    store_p<double>(Wn, m, to_line(w, AX), false);
}
template <int D, int AX>
static void split_axis_t(const StepCtx& c, const Soa& in, const Soa& out, hipStream_t s) {
    if constexpr (AX < D) {
        StepCtx cs = c;
        cs.W = in;
        const Ext& ce = c.ce;
        const int g = ce.g;
        int lo[3] = {0, 0, 0}, hi[3] = {1, 1, 1};
        for (int d = 0; d < D; d++) {
            lo[d] = d == AX ? -g + 1 : -g;
            hi[d] = ce.n[d] + g;
        }
        const Box b = make_box(D, lo[0], hi[0], lo[1], hi[1], lo[2], hi[2]);
        sweep<D, false, RT, false>(cs, AX, b, fs(c.F[AX]), ST_FLUX_X + AX, s);
        hipLaunchKernelGGL((k_apply_axis<D, AX>), dim3(blocks(ce.size)), dim3(256), 0, s, in, out, fs(c.F[AX]), ce, c.dt / c.d[AX], c.ph);
    }
}
void launch_split_axis(const StepCtx& c, int ax, Soa in, Soa out, hipStream_t s) {
    if (c.dim == 2) {
        if (ax == 0) split_axis_t<2, 0>(c, in, out, s);
        else split_axis_t<2, 1>(c, in, out, s);
    } else if (c.dim == 3) {
        if (ax == 0) split_axis_t<3, 0>(c, in, out, s);
        else if (ax == 1) split_axis_t<3, 1>(c, in, out, s);
        else split_axis_t<3, 2>(c, in, out, s);
    }
}

void launch_step(const StepCtx& c, hipStream_t s) {
    if (c.dim == 1) c.mhd ? step_t<1, true, RT>(c, s) : step_t<1, false, RT>(c, s);
    else if (c.dim == 2) c.mhd ? step_t<2, true, RT>(c, s) : step_t<2, false, RT>(c, s);
    else c.mhd ? step_t<3, true, RT>(c, s) : step_t<3, false, RT>(c, s);
}

}  // namespace DG_NS
}  // namespace dg
