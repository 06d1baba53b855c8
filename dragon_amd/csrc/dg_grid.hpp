// dg_grid.hpp -- geometry, SoA layout and failure reporting shared by the
// kernels and the host runtime.
//
// HBM layout: every cell field is its own FP64 array in the reference's
// ExtendedArray order (ghosts included, last index fastest), so a wavefront
// reading consecutive k (3D) or j (2D) is fully coalesced.  Cell records are
// 8 fields (rho, p, vx, vy, vz, Bx, By, Bz; hydro uses the first 5).  Face
// magnetic fields live in three arrays over the (n+1)-sized face/edge index
// space of B (2D: .z is cell-centred, as in the reference).
#pragma once
#include <hip/hip_runtime.h>

namespace dg {

enum { F_RHO = 0, F_P = 1, F_VX = 2, F_VY = 3, F_VZ = 4, F_BX = 5, F_BY = 6, F_BZ = 7 };

// Every array of a grid holds fewer than DG_MAX_ELEMS elements (dg_create
// refuses larger grids), so device code indexes in 32 bits: element offsets
// below 2^31 and byte offsets below 2^32 from a field's base pointer.
constexpr long DG_MAX_ELEMS = 1L << 29;

// Index space of one ExtendedArray (cells: n, faces/edges: n+1).
struct Ext {
    int n[3];
    int g;
    long s[3];    // strides of i, j, k (0 for unused dims)
    long base;    // offset of index (0,0,0)
    long size;    // total elements incl. ghosts
    __host__ __device__ long at(int i, int j, int k) const {
        return (long)((int)base + i * (int)s[0] + j * (int)s[1] + k * (int)s[2]);
    }
};

__host__ __device__ inline Ext make_ext(int dim, int nx, int ny, int nz, int g) {
    Ext e;
    e.n[0] = nx; e.n[1] = dim >= 2 ? ny : 1; e.n[2] = dim >= 3 ? nz : 1;
    e.g = g;
    long ex = nx + 2 * g, ey = dim >= 2 ? ny + 2 * g : 1, ez = dim >= 3 ? nz + 2 * g : 1;
    e.s[2] = dim >= 3 ? 1 : 0;
    e.s[1] = dim >= 2 ? ez : 0;
    e.s[0] = dim >= 3 ? ey * ez : (dim == 2 ? ey : 1);
    e.base = g * (e.s[0] + e.s[1] + e.s[2]);
    e.size = ex * ey * ez;
    return e;
}

// An index box [lo, hi) walked by one thread per point, last dim fastest.
struct Box {
    int lo[3], hi[3];
    __host__ __device__ long count() const {
        return (long)(hi[0] - lo[0]) * (hi[1] - lo[1]) * (hi[2] - lo[2]);
    }
    __device__ bool point(long t, int& i, int& j, int& k) const {
        if (t >= count()) return false;
        const unsigned nk = hi[2] - lo[2], nj = hi[1] - lo[1];
        unsigned r = (unsigned)t;
        k = lo[2] + (int)(r % nk);
        r /= nk;
        j = lo[1] + (int)(r % nj);
        i = lo[0] + (int)(r / nj);
        return true;
    }
};
__host__ __device__ inline Box make_box(int dim, int l0, int h0, int l1 = 0, int h1 = 1, int l2 = 0, int h2 = 1) {
    Box b;
    b.lo[0] = l0; b.hi[0] = h0;
    b.lo[1] = dim >= 2 ? l1 : 0; b.hi[1] = dim >= 2 ? h1 : 1;
    b.lo[2] = dim >= 3 ? l2 : 0; b.hi[2] = dim >= 3 ? h2 : 1;
    return b;
}

// SoA record of nf fields, field f at p + f*stride.
template <class T>
struct SoaT {
    T* p;
    long stride;
    __device__ T& operator()(int f, long m) const {  // the field's base, then a 32-bit byte offset
        return *reinterpret_cast<T*>(reinterpret_cast<char*>(p + f * stride) + (unsigned)((unsigned)m * (unsigned)sizeof(T)));
    }
};
using Soa = SoaT<double>;

// Failure word: the smallest key wins, so the earliest stage and, within a
// stage, the first index in the reference's loop order is what the host
// reports (the reference throws at its first failure).
// key = stage << 56 | sequence << 8 | code
enum { E_RIEMANN_INPUT = 1, E_RIEMANN_SOLUTION = 2, E_NAN = 3, E_UNPHYSICAL = 4 };
enum {
    ST_PRED_X = 1, ST_PRED_Y = 2, ST_PRED_Z = 3,
    ST_CTU_F0 = 4,  // 3D hydro corner fluxes: 4..9
    ST_WHALF_NAN = 10,
    ST_FLUX_X = 11, ST_FLUX_Y = 12, ST_FLUX_Z = 13,
    ST_APPLY_NAN = 14, ST_UNPHYSICAL = 15
};
__device__ __forceinline__ void report(unsigned long long* err, int stage, unsigned long long seq, int code) {
    unsigned long long key = ((unsigned long long)stage << 56) | ((seq & 0xFFFFFFFFFFFFull) << 8) | (unsigned long long)code;
    atomicMin(err, key);
}

}  // namespace dg
