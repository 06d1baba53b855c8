// dg_math.hpp -- device state algebra, limiters, MUSCL-Hancock and the
// Riemann solvers of DRAGON's unsplit Godunov step, for CDNA4.
//
// Everything here runs per face / per cell in registers.  Types are templated
// on the arithmetic type T: T = double is the drop-in (bit-compatible with the
// reference CPU path when built with -ffp-contract=off); T = float is the
// reconstruction/Riemann half of the mixed-precision form.
//
// Operation order follows the reference (cited per function, paths relative
// to DRAGON/), because the FP64 form must round identically.  min/max use the
// std::min/std::max comparison forms so NaN handling matches too.
#pragma once
#include <hip/hip_runtime.h>

namespace dg {

// Physical constants and Config.h switches that the kernels need at run time.
struct Phys {
    double gamma, gm1;          // gamma, gamma - 1.0
    double k1_4pi, k1_8pi, sq4pi;
    // Exact-solver gamma combinations (Constants.h:33-40)
    double Gm1_2G, Gp1_2G, _2G_Gm1, Gm1_2, _2_Gm1, _2_Gp1, Gm1_Gp1, Ginv;
    int limiter;                // Limiters.cpp dispatch
    int muscl;                  // MUSCL_Hancock
    int solver;                 // 0 exact 1 HLL 2 HLLC 3 HLLD 4 HLLE 5 HLLX
    int verify;                 // RIEMANN_VERIFY_FALLBACK
    int try_hlle;               // RIEMANN_FALLBACK_TRY_HLLE
    int hlld_safety;            // HLLD_PHYSICAL_SAFETY
    double fallback_param;      // Riemann_ExactFallback_Parameter
    double exact_tol;
    int exact_iters;
    int exact_rare_check;
    int ct_energy;              // 0 total-E, 1 thermal, 2 beta-gated
    double ct_beta;
    double eightpi;             // 8*M_PI
};

enum { LIM_MINMOD = 0, LIM_MC = 1, LIM_VANLEER = 2, LIM_SUPERBEE = 3, LIM_VANALBADA = 4 };
enum { RS_EXACT = 0, RS_HLL = 1, RS_HLLC = 2, RS_HLLD = 3, RS_HLLE = 4, RS_HLLX = 5 };
// Riemann failure codes (Flux.cpp:94-95)
enum { RF_OK = 0, RF_BAD_INPUT = 1, RF_BAD_SOLUTION = 2 };

template <class T> __device__ __forceinline__ T dmin(T a, T b) { return (b < a) ? b : a; }  // std::min
template <class T> __device__ __forceinline__ T dmax(T a, T b) { return (a < b) ? b : a; }  // std::max
template <class T> __device__ __forceinline__ T dabs(T a) { return __builtin_fabs(a); }
__device__ __forceinline__ double dsqrt(double a) { return __builtin_sqrt(a); }
__device__ __forceinline__ float dsqrt(float a) { return __builtin_sqrtf(a); }
__device__ __forceinline__ bool fin(double a) { return __builtin_isfinite(a); }
__device__ __forceinline__ bool fin(float a) { return __builtin_isfinite(a); }

// Primitive record: rho, p, v, B (FluidElement.hpp:34-43).
template <class T> struct P {
    T rho, p, vx, vy, vz, Bx, By, Bz;
};
// c ? a : b field by field (a select of whole records can end up as a
// private-memory copy addressed by the condition)
template <class T> __device__ __forceinline__ P<T> psel(bool c, const P<T>& a, const P<T>& b) {
    return P<T>{c ? a.rho : b.rho, c ? a.p : b.p, c ? a.vx : b.vx, c ? a.vy : b.vy,
                c ? a.vz : b.vz, c ? a.Bx : b.Bx, c ? a.By : b.By, c ? a.Bz : b.Bz};
}
// Conservative record / flux: rho, E, mom, B (FluidElement.hpp:70-80).
template <class T> struct U {
    T rho, E, mx, my, mz, Bx, By, Bz;
};

template <class T> __device__ __forceinline__ P<T> pzero() { return P<T>{0, 0, 0, 0, 0, 0, 0, 0}; }
template <class T> __device__ __forceinline__ U<T> uzero() { return U<T>{0, 0, 0, 0, 0, 0, 0, 0}; }

template <class T> __device__ __forceinline__ T dot3(T ax, T ay, T az, T bx, T by, T bz) { return ax * bx + ay * by + az * bz; }

// ---- U arithmetic (Arithmetic.cpp:14-128), hydro leaves B untouched
template <bool MHD, class T> __device__ __forceinline__ U<T> uadd(U<T> a, const U<T>& b) {
    a.rho += b.rho; a.mx += b.mx; a.my += b.my; a.mz += b.mz; a.E += b.E;
    if constexpr (MHD) { a.Bx += b.Bx; a.By += b.By; a.Bz += b.Bz; }
    return a;
}
template <bool MHD, class T> __device__ __forceinline__ U<T> usub(U<T> a, const U<T>& b) {
    a.rho -= b.rho; a.mx -= b.mx; a.my -= b.my; a.mz -= b.mz; a.E -= b.E;
    if constexpr (MHD) { a.Bx -= b.Bx; a.By -= b.By; a.Bz -= b.Bz; }
    return a;
}
template <bool MHD, class T> __device__ __forceinline__ U<T> umul(U<T> a, T s) {
    a.rho *= s; a.mx *= s; a.my *= s; a.mz *= s; a.E *= s;
    if constexpr (MHD) { a.Bx *= s; a.By *= s; a.Bz *= s; }
    return a;
}
template <bool MHD, class T> __device__ __forceinline__ U<T> udiv(U<T> a, T s) {
    a.rho /= s; a.mx /= s; a.my /= s; a.mz /= s; a.E /= s;
    if constexpr (MHD) { a.Bx /= s; a.By /= s; a.Bz /= s; }
    return a;
}

// ---- conversions (FluidElement.cpp:35-66)
template <bool MHD, class T> __device__ __forceinline__ T energy(const P<T>& w, const Phys& ph) {
    T E = w.p / T(ph.gm1) + (w.rho / 2) * dot3(w.vx, w.vy, w.vz, w.vx, w.vy, w.vz);
    if constexpr (MHD) E += dot3(w.Bx, w.By, w.Bz, w.Bx, w.By, w.Bz) * T(ph.k1_8pi);
    return E;
}
template <bool MHD, class T> __device__ __forceinline__ T pressure(const U<T>& u, const Phys& ph) {
    T Et = u.E - dot3(u.mx, u.my, u.mz, u.mx, u.my, u.mz) / (2 * u.rho);
    if constexpr (MHD) Et -= dot3(u.Bx, u.By, u.Bz, u.Bx, u.By, u.Bz) * T(ph.k1_8pi);
    return T(ph.gm1) * Et;
}
template <bool MHD, class T> __device__ __forceinline__ U<T> cons(const P<T>& w, const Phys& ph) {
    U<T> u;
    u.rho = w.rho;
    u.mx = w.vx * w.rho; u.my = w.vy * w.rho; u.mz = w.vz * w.rho;
    u.E = energy<MHD>(w, ph);
    if constexpr (MHD) { u.Bx = w.Bx; u.By = w.By; u.Bz = w.Bz; } else { u.Bx = u.By = u.Bz = 0; }
    return u;
}
template <bool MHD, class T> __device__ __forceinline__ P<T> prim(const U<T>& u, const Phys& ph) {
    P<T> w;
    w.rho = u.rho;
    w.vx = u.mx / u.rho; w.vy = u.my / u.rho; w.vz = u.mz / u.rho;
    w.p = pressure<MHD>(u, ph);
    if constexpr (MHD) { w.Bx = u.Bx; w.By = u.By; w.Bz = u.Bz; } else { w.Bx = w.By = w.Bz = 0; }
    return w;
}
template <bool MHD, class T> __device__ __forceinline__ T enthalpy(const P<T>& w, const Phys& ph) {  // :69-75
    T h = energy<MHD>(w, ph) + w.p;
    if constexpr (MHD) h += dot3(w.Bx, w.By, w.Bz, w.Bx, w.By, w.Bz) * T(ph.k1_8pi);
    return h / w.rho;
}
template <bool MHD, class T> __device__ __forceinline__ bool physical(const P<T>& w) {  // :82-87
    if constexpr (MHD) { if (!(fin(w.Bx) && fin(w.By) && fin(w.Bz))) return false; }
    return fin(w.rho) && fin(w.vx) && fin(w.vy) && fin(w.vz) && fin(w.p) && w.rho > T(0) && w.p > T(0);
}
template <bool MHD, class T> __device__ __forceinline__ bool finite(const U<T>& u) {  // :88-93
    if constexpr (MHD) { if (!(fin(u.Bx) && fin(u.By) && fin(u.Bz))) return false; }
    return fin(u.rho) && fin(u.mx) && fin(u.my) && fin(u.mz) && fin(u.E);
}
template <bool MHD, class T> __device__ __forceinline__ bool physical(const U<T>& u, const Phys& ph) {  // :94-96
    return finite<MHD>(u) && u.rho > T(0) && pressure<MHD>(u, ph) > T(0);
}

// ---- wave speeds (FluidElement.cpp:99-120)
template <class T> __device__ __forceinline__ T cs(const P<T>& w, const Phys& ph) { return dsqrt(T(ph.gamma) * w.p / w.rho); }
template <class T> __device__ __forceinline__ T c_fast(const P<T>& w, T Bk, const Phys& ph) {
    T c = cs(w, ph);
    T a = dsqrt(dot3(w.Bx, w.By, w.Bz, w.Bx, w.By, w.Bz) / w.rho) / T(ph.sq4pi);
    T c2 = c * c, a2 = a * a, ak2 = (Bk * Bk / w.rho) * T(ph.k1_4pi);
    T s = c2 + a2;
    T disc = dmax(T(0), (s * s) * T(0.25) - ak2 * c2);
    return dsqrt((c2 + a2) * T(0.5) + dsqrt(disc));
}
template <class T> __device__ __forceinline__ T c_fast_max(const P<T>& w, const Phys& ph) {
    T Bk;
    if (w.Bx * w.Bx < w.By * w.By) Bk = (w.Bx * w.Bx < w.Bz * w.Bz) ? w.Bx : w.Bz;
    else Bk = (w.By * w.By < w.Bz * w.Bz) ? w.By : w.Bz;
    return c_fast(w, Bk, ph);
}

// ---- x-direction physical flux (FluidElement.cpp:133-155)
template <bool MHD, class T> __device__ __forceinline__ U<T> flux(const U<T>& u, T vx, T vy, T vz, const Phys& ph) {
    U<T> F;
    T p = pressure<MHD>(u, ph);
    F.rho = u.mx;
    F.mx = u.mx * vx; F.my = u.my * vx; F.mz = u.mz * vx;
    F.mx += p;
    F.E = (u.E + p) * vx;
    if constexpr (MHD) {
        T pm = dot3(u.Bx, u.By, u.Bz, u.Bx, u.By, u.Bz) * T(ph.k1_8pi);
        F.mx += pm;
        F.E += pm * vx;
        T tx = (u.Bx * u.Bx) * T(ph.k1_4pi), ty = (u.By * u.Bx) * T(ph.k1_4pi), tz = (u.Bz * u.Bx) * T(ph.k1_4pi);
        F.mx -= tx; F.my -= ty; F.mz -= tz;
        F.E -= dot3(vx, vy, vz, tx, ty, tz);
        F.By = u.By * vx - vy * u.Bx;
        F.Bz = u.Bz * vx - vz * u.Bx;
        F.Bx = 0;
    } else {
        F.Bx = F.By = F.Bz = 0;
    }
    return F;
}
template <bool MHD, class T> __device__ __forceinline__ U<T> pflux(const P<T>& w, const Phys& ph) {  // :124
    return flux<MHD>(cons<MHD>(w, ph), w.vx, w.vy, w.vz, ph);
}

template <class T> __device__ __forceinline__ P<T> swapXY(P<T> w) { T t = w.vx; w.vx = w.vy; w.vy = t; t = w.Bx; w.Bx = w.By; w.By = t; return w; }
template <class T> __device__ __forceinline__ P<T> swapXZ(P<T> w) { T t = w.vx; w.vx = w.vz; w.vz = t; t = w.Bx; w.Bx = w.Bz; w.Bz = t; return w; }
template <class T> __device__ __forceinline__ U<T> swapXY(U<T> u) { T t = u.mx; u.mx = u.my; u.my = t; t = u.Bx; u.Bx = u.By; u.By = t; return u; }
template <class T> __device__ __forceinline__ U<T> swapXZ(U<T> u) { T t = u.mx; u.mx = u.mz; u.mz = t; t = u.Bx; u.Bx = u.Bz; u.Bz = t; return u; }

// ---- limiters (Limiters.cpp:88-191)
template <class T> __device__ __forceinline__ T minmod(T a, T b) {
    if (a * b <= T(0)) return T(0);
    return (dabs(a) < dabs(b)) ? a : b;
}
template <class T> __device__ __forceinline__ T limit1(T a, T b, int lim) {
    if (a * b <= T(0)) return T(0);
    switch (lim) {
        case LIM_MC: {
            T c = dabs(a + b);
            if (c < 4 * dabs(a) && c < 4 * dabs(b)) return a < T(0) ? -c / 2 : c / 2;
            return 2 * minmod(a, b);
        }
        case LIM_VANLEER: return 2 * a * b / (a + b);
        case LIM_SUPERBEE: {
            T A = dabs(a), Bb = dabs(b);
            T m = dmax(dmin(2 * A, Bb), dmin(A, 2 * Bb));
            return a < T(0) ? -m : m;
        }
        case LIM_VANALBADA: return a * b * (a + b) / (a * a + b * b);
        default: return (dabs(a) < dabs(b)) ? a : b;
    }
}

// ---- MUSCL-Hancock (MUSCL.cpp:16-61).  dB = (dBx, dBy, dBz) in the swapped
// frame; falls back to first order if either side is unphysical.
template <bool MHD, class T>
__device__ __forceinline__ void muscl(const P<T>& wL, P<T>& L, const P<T>& wC, P<T>& R, const P<T>& wR, T dt_dL,
                                      T dBx, T dBy, T dBz, const Phys& ph) {
    // The result is selected field by field: assigning whole records on
    // the two paths left a private-memory copy in one instantiation.
    bool ok = false;
    P<T> l = wC, r = wC;
    if (ph.muscl) {
        const int lim = ph.limiter;
        P<T> dW;
        dW.rho = limit1(wC.rho - wL.rho, wR.rho - wC.rho, lim);
        dW.vx = limit1(wC.vx - wL.vx, wR.vx - wC.vx, lim);
        dW.vy = limit1(wC.vy - wL.vy, wR.vy - wC.vy, lim);
        dW.vz = limit1(wC.vz - wL.vz, wR.vz - wC.vz, lim);
        dW.p = limit1(wC.p - wL.p, wR.p - wC.p, lim);
        if constexpr (MHD) {
            dW.Bx = limit1(wC.Bx - wL.Bx, wR.Bx - wC.Bx, lim);
            dW.By = limit1(wC.By - wL.By, wR.By - wC.By, lim);
            dW.Bz = limit1(wC.Bz - wL.Bz, wR.Bz - wC.Bz, lim);
        }
        l.rho = wC.rho - T(0.5) * dW.rho; r.rho = wC.rho + T(0.5) * dW.rho;
        l.vx = wC.vx - dW.vx * T(0.5); r.vx = wC.vx + dW.vx * T(0.5);
        l.vy = wC.vy - dW.vy * T(0.5); r.vy = wC.vy + dW.vy * T(0.5);
        l.vz = wC.vz - dW.vz * T(0.5); r.vz = wC.vz + dW.vz * T(0.5);
        l.p = wC.p - T(0.5) * dW.p; r.p = wC.p + T(0.5) * dW.p;
        if constexpr (MHD) {
            l.Bx = wC.Bx - dW.Bx * T(0.5); r.Bx = wC.Bx + dW.Bx * T(0.5);
            l.By = wC.By - dW.By * T(0.5); r.By = wC.By + dW.By * T(0.5);
            l.Bz = wC.Bz - dW.Bz * T(0.5); r.Bz = wC.Bz + dW.Bz * T(0.5);
        } else {
            l.Bx = l.By = l.Bz = r.Bx = r.By = r.Bz = 0;
        }
        U<T> UL = cons<MHD>(l, ph), UR = cons<MHD>(r, ph);
        U<T> corr = umul<MHD>(usub<MHD>(flux<MHD>(UR, r.vx, r.vy, r.vz, ph), flux<MHD>(UL, l.vx, l.vy, l.vz, ph)), T(0.5) * dt_dL);
        if constexpr (MHD) {
            const T a_s = (r.Bx - l.Bx) * dt_dL;
            const T vy = T(0.5) * (l.vy + r.vy), vz = T(0.5) * (l.vz + r.vz);
            const T dBy_ = T(0.5) * wC.vy * minmod(dBx, -dBy) - T(0.5) * vy * a_s;
            const T dBz_ = T(0.5) * wC.vz * minmod(dBx, -dBz) - T(0.5) * vz * a_s;
            corr.By -= dBy_;
            corr.Bz -= dBz_;
            corr.E -= T(ph.k1_4pi) * (wC.By * dBy_ + wC.Bz * dBz_);
        }
        l = prim<MHD>(usub<MHD>(UL, corr), ph);
        r = prim<MHD>(usub<MHD>(UR, corr), ph);
        ok = physical<MHD>(l) && physical<MHD>(r);
    }
    L = psel(ok, l, wC);
    R = psel(ok, r, wC);
}

// ---- HLL family (HLL.cpp)
template <bool MHD, class T> __device__ __forceinline__ U<T> hll_s(const P<T>& L, const P<T>& R, T sl, T sr, const Phys& ph) {
    if (sl >= T(0)) return pflux<MHD>(L, ph);
    if (sr <= T(0)) return pflux<MHD>(R, ph);
    U<T> UL = cons<MHD>(L, ph), UR = cons<MHD>(R, ph);
    U<T> FL = flux<MHD>(UL, L.vx, L.vy, L.vz, ph), FR = flux<MHD>(UR, R.vx, R.vy, R.vz, ph);
    return udiv<MHD>(uadd<MHD>(usub<MHD>(umul<MHD>(FL, sr), umul<MHD>(FR, sl)), umul<MHD>(usub<MHD>(UR, UL), sl * sr)), sr - sl);
}
// Roe average (HLL.cpp:21-37)
template <bool MHD, class T> __device__ __forceinline__ P<T> roe_avg(const P<T>& L, const P<T>& R, const Phys& ph) {
    T sql = dsqrt(L.rho), sqr = dsqrt(R.rho);
    T a = sql / (sql + sqr), b = sqr / (sql + sqr);
    P<T> w;
    w.rho = sql * sqr;
    w.vx = L.vx * a + R.vx * b; w.vy = L.vy * a + R.vy * b; w.vz = L.vz * a + R.vz * b;
    if constexpr (MHD) { w.Bx = L.Bx * a + R.Bx * b; w.By = L.By * a + R.By * b; w.Bz = L.Bz * a + R.Bz * b; }
    else { w.Bx = w.By = w.Bz = 0; }
    T H = a * enthalpy<MHD>(L, ph) + b * enthalpy<MHD>(R, ph);
    T a2 = T(ph.gm1) * (H - dot3(w.vx, w.vy, w.vz, w.vx, w.vy, w.vz) / 2);
    w.p = a2 * w.rho / T(ph.gamma);
    return w;
}
template <bool MHD, class T> __device__ U<T> hll(P<T> L, P<T> R, const Phys& ph) {  // HLL.cpp:42-60
    P<T> M = roe_avg<MHD>(L, R, ph);
    T aL, aR, aM;
    if constexpr (MHD) {
        T Bx = (L.Bx + R.Bx) / 2;
        L.Bx = Bx; R.Bx = Bx;
        aL = c_fast(L, L.Bx, ph); aR = c_fast(R, R.Bx, ph); aM = c_fast(M, M.Bx, ph);
    } else {
        aL = cs(L, ph); aR = cs(R, ph); aM = cs(M, ph);
    }
    return hll_s<MHD>(L, R, dmin(L.vx - aL, M.vx - aM), dmax(R.vx + aR, M.vx + aM), ph);
}
template <bool MHD, class T> __device__ U<T> hlle(P<T> L, P<T> R, const Phys& ph) {  // HLL.cpp:61-83
    T sql = dsqrt(L.rho), sqr = dsqrt(R.rho);
    T a = sql / (sql + sqr), b = sqr / (sql + sqr);
    T aL, aR;
    if constexpr (MHD) {
        T Bx = (L.Bx + R.Bx) / 2;
        L.Bx = Bx; R.Bx = Bx;
        aL = c_fast(L, L.Bx, ph); aR = c_fast(R, R.Bx, ph);
    } else {
        aL = cs(L, ph); aR = cs(R, ph);
    }
    T v = (R.vx - L.vx) / (sql + sqr);
    T d = (a * aL * aL + b * aR * aR) + T(0.5) * (sql * sqr) * v * v;
    d = dsqrt(d);
    T u = a * L.vx + b * R.vx;
    return hll_s<MHD>(L, R, dmin(L.vx - aL, u - d), dmax(R.vx + aR, u + d), ph);
}
// HLLC (HLL.cpp:103-132), hydro
template <class T> __device__ U<T> hllc(const P<T>& L, const P<T>& R, const Phys& ph) {
    P<T> M = roe_avg<false>(L, R, ph);
    T aL = cs(L, ph), aR = cs(R, ph), aM = cs(M, ph);
    T sl = dmin(L.vx - aL, M.vx - aM);
    T sr = dmax(R.vx + aR, M.vx + aM);
    if (sl >= T(0)) return pflux<false>(L, ph);
    if (sr <= T(0)) return pflux<false>(R, ph);
    T pr = R.p + R.rho * R.vx * (R.vx - sr), pl = L.p + L.rho * L.vx * (L.vx - sl);
    T sc = (pr - pl) / (R.rho * (R.vx - sr) - L.rho * (L.vx - sl));
    const P<T> X = psel(sc > T(0), L, R);
    T sx = sc > T(0) ? sl : sr;
    U<T> UX = cons<false>(X, ph);
    U<T> Us = udiv<false>(umul<false>(UX, sx - X.vx), sx - sc);
    Us.mx = Us.rho * sc;
    Us.E += Us.rho * (sc - X.vx) * (sc + X.p / (X.rho * (sx - X.vx)));
    return uadd<false>(flux<false>(UX, X.vx, X.vy, X.vz, ph), umul<false>(usub<false>(Us, UX), sx));
}

// ---- HLLD (HLLD.cpp), MHD
template <class T>
__device__ __forceinline__ void outer_star(U<T>& us, const P<T>& K, T xK, T SK, T SM, T pT, T Bx, const Phys& ph) {
    T Bx2 = (Bx * Bx) * T(ph.k1_4pi);  // HLLD.cpp:21-37
    T den = xK * (SK - SM) - Bx2;
    bool tiny = dabs(den) < T(1e-12);
    if (tiny) { us.Bx = K.Bx; us.By = K.By; us.Bz = K.Bz; }
    else {
        T s = xK * (SK - K.vx) - Bx2;
        us.Bx = K.Bx * s / den; us.By = K.By * s / den; us.Bz = K.Bz * s / den;
    }
    us.Bx = Bx;
    T vsx, vsy, vsz;
    if (tiny) { vsx = K.vx - T(0); vsy = K.vy - T(0); vsz = K.vz - T(0); }
    else {
        T s = SM - K.vx;
        vsx = K.vx - K.Bx * Bx * s / den * T(ph.k1_4pi);
        vsy = K.vy - K.By * Bx * s / den * T(ph.k1_4pi);
        vsz = K.vz - K.Bz * Bx * s / den * T(ph.k1_4pi);
    }
    vsx = SM;
    us.mx = vsx * us.rho; us.my = vsy * us.rho; us.mz = vsz * us.rho;
    us.E = energy<true>(K, ph) * (SK - K.vx);
    us.E += pT * SM - (K.p + dot3(K.Bx, K.By, K.Bz, K.Bx, K.By, K.Bz) * T(ph.k1_8pi)) * K.vx;
    us.E += Bx * (dot3(K.vx, K.vy, K.vz, K.Bx, K.By, K.Bz) - dot3(vsx, vsy, vsz, us.Bx, us.By, us.Bz)) * T(ph.k1_4pi);
    us.E /= (SK - SM);
}
template <class T>
__device__ __forceinline__ void outer_star_zero(U<T>& us, const P<T>& K, T SK, T SM, T pT, const Phys& ph) {  // :137-150
    if (dabs(SK - SM) < T(1e-12)) { us.Bx = K.Bx; us.By = K.By; us.Bz = K.Bz; }
    else { T s = SK - K.vx, d = SK - SM; us.Bx = K.Bx * s / d; us.By = K.By * s / d; us.Bz = K.Bz * s / d; }
    us.Bx = 0;
    us.mx = SM * us.rho; us.my = K.vy * us.rho; us.mz = K.vz * us.rho;
    us.E = energy<true>(K, ph) * (SK - K.vx);
    us.E += pT * SM - (K.p + dot3(K.Bx, K.By, K.Bz, K.Bx, K.By, K.Bz) * T(ph.k1_8pi)) * K.vx;
    us.E /= (SK - SM);
}
template <class T> __device__ __forceinline__ P<T> star_prim(const U<T>& us, T pT, const Phys& ph) {
    P<T> w;
    w.rho = us.rho;
    w.vx = us.mx / us.rho; w.vy = us.my / us.rho; w.vz = us.mz / us.rho;
    w.Bx = us.Bx; w.By = us.By; w.Bz = us.Bz;
    w.p = pT - dot3(w.Bx, w.By, w.Bz, w.Bx, w.By, w.Bz) * T(ph.k1_8pi);
    return w;
}
// Floor of the zero-normal-B test's scale (HLLD.cpp:50-51 uses 1e-100).
// 1e-100 is zero in FP32, which would send faces with B = 0 exactly past the
// zero-B solver; the FP32 form uses 1e-30 so they take the same branch.
template <class T> __device__ __forceinline__ T hlld_floor() { return sizeof(T) == 4 ? T(1e-30) : T(1e-100); }
template <bool MHD, class T> __device__ __forceinline__ U<T> usel(bool c, const U<T>& a, const U<T>& b) {
    U<T> r;
    r.rho = c ? a.rho : b.rho; r.E = c ? a.E : b.E; r.mx = c ? a.mx : b.mx; r.my = c ? a.my : b.my; r.mz = c ? a.mz : b.mz;
    if constexpr (MHD) { r.Bx = c ? a.Bx : b.Bx; r.By = c ? a.By : b.By; r.Bz = c ? a.Bz : b.Bz; }
    else { r.Bx = r.By = r.Bz = T(0); }
    return r;
}
// Outer-star flux F_K + S_K (U*_K - U_K) with the normal B flux zeroed.
template <class T> __device__ __forceinline__ U<T> star_flux(const P<T>& K, const U<T>& usK, T SK, const Phys& ph) {
    U<T> f = uadd<true>(pflux<true>(K, ph), umul<true>(usub<true>(usK, cons<true>(K, ph)), SK));
    f.Bx = 0;
    return f;
}
// HLLD for a wavefront.  The reference's solver (HLLD.cpp) branches on the
// wave region of the face: supersonic left/right, outer star left/right,
// double star.  In a wavefront whose faces fall in different regions every
// branch would run in turn.  Here every face of the nonzero-normal-B solver
// follows one instruction stream: both outer star states and the double-star
// state are built, the upwind side K is chosen per face, and the result is
// selected from F(K), F*K = F(K) + S_K (U*K - U_K) and F*K + S*K (U**K - U*K).
// Each face returns the value the reference's branch for its region returns,
// computed by the same operations in the same order; values of the regions a
// face is not in are discarded by the selects (never multiplied by zero).
// The physical-state safety (HLLD_PHYSICAL_SAFETY -> HLLE) is applied under the
// reference's conditions, and the zero-normal-B solver (HLLD.cpp:137-205)
// keeps its own branch.
template <class T> __device__ U<T> hlld_uniform(P<T> L, P<T> R, const Phys& ph) {  // HLLD.cpp:38-133, :152-205
    T Bx = (L.Bx + R.Bx) / 2;
    L.Bx = Bx; R.Bx = Bx;
    T cl = c_fast(L, L.Bx, ph), cr = c_fast(R, R.Bx, ph);
    T SL = dmin(L.vx - cl, R.vx - cr);
    T SR = dmax(L.vx + cl, R.vx + cr);
    const bool supL = SL >= T(0), supR = !supL && SR <= T(0);
    T cutoff = dmax(dot3(L.Bx, L.By, L.Bz, L.Bx, L.By, L.Bz) + dot3(R.Bx, R.By, R.Bz, R.Bx, R.By, R.Bz), hlld_floor<T>());
    const bool zeroB = Bx * Bx < T(1e-12) * cutoff;
    T pTL = L.p + dot3(L.Bx, L.By, L.Bz, L.Bx, L.By, L.Bz) * T(ph.k1_8pi);
    T pTR = R.p + dot3(R.Bx, R.By, R.Bz, R.Bx, R.By, R.Bz) * T(ph.k1_8pi);
    T xL = L.rho * (SL - L.vx), xR = R.rho * (SR - R.vx);
    T SM = ((xR * R.vx - xL * L.vx) - (pTR - pTL)) / (xR - xL);
    U<T> usL = uzero<T>(), usR = uzero<T>();
    usL.rho = xL / (SL - SM);
    usR.rho = xR / (SR - SM);
    T pT = (xR * pTL - xL * pTR + xL * xR * (R.vx - L.vx)) / (xR - xL);
    P<T> wsL, wsR;
    if (zeroB && !supL && !supR) {  // HLLD_zero_B (HLLD.cpp:152-205)
        wsL = pzero<T>(); wsR = pzero<T>();
        if (SM >= T(0)) {
            outer_star_zero(usL, L, SL, SM, pT, ph);
            wsL = star_prim(usL, pT, ph);
            if (ph.hlld_safety && (!physical<true>(wsL) || !finite<true>(usL))) return hlle<true>(L, R, ph);
        }
        if (SM <= T(0)) {
            outer_star_zero(usR, R, SR, SM, pT, ph);
            wsR = star_prim(usR, pT, ph);
            if (ph.hlld_safety && (!physical<true>(wsR) || !finite<true>(usR))) return hlle<true>(L, R, ph);
        }
        if (SM > T(0)) return star_flux(L, usL, SL, ph);
        if (SM < T(0)) return star_flux(R, usR, SR, ph);
        T sql = dsqrt(wsL.rho), sqr = dsqrt(wsR.rho);
        T a = sql / (sql + sqr), b = sqr / (sql + sqr);
        return uadd<true>(umul<true>(star_flux(L, usL, SL, ph), a), umul<true>(star_flux(R, usR, SR, ph), b));
    }
    T SsL = SM - dabs(Bx) / (T(ph.sq4pi) * dsqrt(usL.rho));
    T SsR = SM + dabs(Bx) / (T(ph.sq4pi) * dsqrt(usR.rho));
    outer_star(usL, L, xL, SL, SM, pT, Bx, ph);
    wsL = star_prim(usL, pT, ph);
    outer_star(usR, R, xR, SR, SM, pT, Bx, ph);
    wsR = star_prim(usR, pT, ph);
    // region of the face, in the reference's order of tests
    const bool sub = !supL && !supR;
    const bool starL = SsL >= T(0), starR = !starL && SsR <= T(0);
    const bool dstar = sub && !starL && !starR;
    const bool useL = supL || (sub && (starL || (!starR && SM >= T(0))));
    bool bad = false;
    if (ph.hlld_safety && sub) {
        bad = (SsR >= T(0) && (!physical<true>(wsL) || !physical<true>(usL, ph))) ||
              (SsL <= T(0) && (!physical<true>(wsR) || !physical<true>(usR, ph)));
    }
    // double-star state (used by dstar faces only)
    const P<T> ws = psel(SM >= T(0), wsL, wsR);
    T sqL = dsqrt(wsL.rho), sqR = dsqrt(wsR.rho);
    T a = sqL / (sqL + sqR), b = sqR / (sqL + sqR);
    T sbx = Bx > T(0) ? T(ph.sq4pi) : (Bx < T(0) ? -T(ph.sq4pi) : T(0));
    T sum = sqL + sqR;
    P<T> wss;
    U<T> uss;
    wss.rho = ws.rho;
    uss.rho = wss.rho;
    wss.vx = SM;
    wss.vy = wsL.vy * a + wsR.vy * b + (usR.By - usL.By) * sbx * T(ph.k1_4pi) / sum;
    wss.vz = wsL.vz * a + wsR.vz * b + (usR.Bz - usL.Bz) * sbx * T(ph.k1_4pi) / sum;
    uss.mx = wss.vx * uss.rho; uss.my = wss.vy * uss.rho; uss.mz = wss.vz * uss.rho;
    wss.Bx = Bx;
    wss.By = usR.By * a + usL.By * b + (wsR.vy - wsL.vy) * sbx * sqL * sqR / sum;
    wss.Bz = usR.Bz * a + usL.Bz * b + (wsR.vz - wsL.vz) * sbx * sqL * sqR / sum;
    uss.Bx = wss.Bx; uss.By = wss.By; uss.Bz = wss.Bz;
    T dE = sbx * (dot3(ws.vx, ws.vy, ws.vz, ws.Bx, ws.By, ws.Bz) - dot3(wss.vx, wss.vy, wss.vz, wss.Bx, wss.By, wss.Bz)) *
           T(ph.k1_4pi) * dsqrt(wss.rho);
    uss.E = SM >= T(0) ? usL.E - dE : usR.E + dE;
    if (ph.hlld_safety && dstar) {
        wss.p = pT - dot3(wss.Bx, wss.By, wss.Bz, wss.Bx, wss.By, wss.Bz) * T(ph.k1_8pi);
        bad = bad || !physical<true>(wss) || !finite<true>(uss);
    }
    // the upwind side's fluxes
    const P<T> K = psel(useL, L, R);
    const U<T> usK = usel<true>(useL, usL, usR);
    const T SK = useL ? SL : SR, SsK = useL ? SsL : SsR;
    const U<T> FK = pflux<true>(K, ph);
    U<T> FsK = uadd<true>(FK, umul<true>(usub<true>(usK, cons<true>(K, ph)), SK));
    FsK.Bx = 0;
    const U<T> FssK = uadd<true>(FsK, umul<true>(usub<true>(uss, usK), SsK));
    U<T> F = usel<true>(supL || supR, FK, usel<true>(dstar, FssK, FsK));
    if (bad) F = hlle<true>(L, R, ph);
    return F;
}

// HLLD in the reference's control flow: one branch per wave region, the
// outer-star fluxes formed where they are returned.
template <class T> __device__ U<T> hlld_branch(P<T> L, P<T> R, const Phys& ph) {  // HLLD.cpp:38-133, :152-205
    T Bx = (L.Bx + R.Bx) / 2;
    L.Bx = Bx; R.Bx = Bx;
    T cl = c_fast(L, L.Bx, ph), cr = c_fast(R, R.Bx, ph);
    T SL = dmin(L.vx - cl, R.vx - cr);
    if (SL >= T(0)) return pflux<true>(L, ph);
    T SR = dmax(L.vx + cl, R.vx + cr);
    if (SR <= T(0)) return pflux<true>(R, ph);
    T cutoff = dmax(dot3(L.Bx, L.By, L.Bz, L.Bx, L.By, L.Bz) + dot3(R.Bx, R.By, R.Bz, R.Bx, R.By, R.Bz), hlld_floor<T>());
    const bool zeroB = Bx * Bx < T(1e-12) * cutoff;
    T pTL = L.p + dot3(L.Bx, L.By, L.Bz, L.Bx, L.By, L.Bz) * T(ph.k1_8pi);
    T pTR = R.p + dot3(R.Bx, R.By, R.Bz, R.Bx, R.By, R.Bz) * T(ph.k1_8pi);
    T xL = L.rho * (SL - L.vx), xR = R.rho * (SR - R.vx);
    T SM = ((xR * R.vx - xL * L.vx) - (pTR - pTL)) / (xR - xL);
    U<T> usL = uzero<T>(), usR = uzero<T>();
    usL.rho = xL / (SL - SM);
    usR.rho = xR / (SR - SM);
    T pT = (xR * pTL - xL * pTR + xL * xR * (R.vx - L.vx)) / (xR - xL);
    P<T> wsL = pzero<T>(), wsR = pzero<T>();
    if (zeroB) {
        if (SM >= T(0)) {
            outer_star_zero(usL, L, SL, SM, pT, ph);
            wsL = star_prim(usL, pT, ph);
            if (ph.hlld_safety && (!physical<true>(wsL) || !finite<true>(usL))) return hlle<true>(L, R, ph);
        }
        if (SM <= T(0)) {
            outer_star_zero(usR, R, SR, SM, pT, ph);
            wsR = star_prim(usR, pT, ph);
            if (ph.hlld_safety && (!physical<true>(wsR) || !finite<true>(usR))) return hlle<true>(L, R, ph);
        }
        if (SM > T(0)) return star_flux(L, usL, SL, ph);
        if (SM < T(0)) return star_flux(R, usR, SR, ph);
        T sql = dsqrt(wsL.rho), sqr = dsqrt(wsR.rho);
        T a = sql / (sql + sqr), b = sqr / (sql + sqr);
        return uadd<true>(umul<true>(star_flux(L, usL, SL, ph), a), umul<true>(star_flux(R, usR, SR, ph), b));
    }
    T SsL = SM - dabs(Bx) / (T(ph.sq4pi) * dsqrt(usL.rho));
    T SsR = SM + dabs(Bx) / (T(ph.sq4pi) * dsqrt(usR.rho));
    if (SsR >= T(0)) {
        outer_star(usL, L, xL, SL, SM, pT, Bx, ph);
        wsL = star_prim(usL, pT, ph);
        if (ph.hlld_safety && (!physical<true>(wsL) || !physical<true>(usL, ph))) return hlle<true>(L, R, ph);
    }
    if (SsL <= T(0)) {
        outer_star(usR, R, xR, SR, SM, pT, Bx, ph);
        wsR = star_prim(usR, pT, ph);
        if (ph.hlld_safety && (!physical<true>(wsR) || !physical<true>(usR, ph))) return hlle<true>(L, R, ph);
    }
    if (SsL >= T(0)) return star_flux(L, usL, SL, ph);
    if (SsR <= T(0)) return star_flux(R, usR, SR, ph);
    const P<T> ws = psel(SM >= T(0), wsL, wsR);
    T sqL = dsqrt(wsL.rho), sqR = dsqrt(wsR.rho);
    T a = sqL / (sqL + sqR), b = sqR / (sqL + sqR);
    T sbx = Bx > T(0) ? T(ph.sq4pi) : (Bx < T(0) ? -T(ph.sq4pi) : T(0));
    T sum = sqL + sqR;
    P<T> wss;
    U<T> uss;
    wss.rho = ws.rho;
    uss.rho = wss.rho;
    wss.vx = SM;
    wss.vy = wsL.vy * a + wsR.vy * b + (usR.By - usL.By) * sbx * T(ph.k1_4pi) / sum;
    wss.vz = wsL.vz * a + wsR.vz * b + (usR.Bz - usL.Bz) * sbx * T(ph.k1_4pi) / sum;
    uss.mx = wss.vx * uss.rho; uss.my = wss.vy * uss.rho; uss.mz = wss.vz * uss.rho;
    wss.Bx = Bx;
    wss.By = usR.By * a + usL.By * b + (wsR.vy - wsL.vy) * sbx * sqL * sqR / sum;
    wss.Bz = usR.Bz * a + usL.Bz * b + (wsR.vz - wsL.vz) * sbx * sqL * sqR / sum;
    uss.Bx = wss.Bx; uss.By = wss.By; uss.Bz = wss.Bz;
    T dE = sbx * (dot3(ws.vx, ws.vy, ws.vz, ws.Bx, ws.By, ws.Bz) - dot3(wss.vx, wss.vy, wss.vz, wss.Bx, wss.By, wss.Bz)) *
           T(ph.k1_4pi) * dsqrt(wss.rho);
    uss.E = SM >= T(0) ? usL.E - dE : usR.E + dE;
    if (ph.hlld_safety) {
        wss.p = pT - dot3(wss.Bx, wss.By, wss.Bz, wss.Bx, wss.By, wss.Bz) * T(ph.k1_8pi);
        if (!physical<true>(wss) || !finite<true>(uss)) return hlle<true>(L, R, ph);
    }
    if (SM >= T(0)) return uadd<true>(star_flux(L, usL, SL, ph), umul<true>(usub<true>(uss, usL), SsL));
    return uadd<true>(star_flux(R, usR, SR, ph), umul<true>(usub<true>(uss, usR), SsR));
}

// FP32 (mixed form) runs the wavefront-uniform solver, FP64 the branching one:
// measured on MI355X at 512^3, the uniform stream saves time in FP32 and
// costs time in FP64 (both outer stars always built).  Both return the
// same bits for every face.
template <class T> __device__ __forceinline__ U<T> hlld(const P<T>& L, const P<T>& R, const Phys& ph) {
    if constexpr (sizeof(T) == 4) return hlld_uniform<T>(L, R, ph);
    else return hlld_branch<T>(L, R, ph);
}

// ---- exact solver (Exact.cpp), hydro fallback only
template <class T> __device__ __forceinline__ T ex_f(T p, const P<T>& w, const Phys& ph) {
    if (p > w.p) {
        T A = T(ph._2_Gp1) / w.rho, Bc = T(ph.Gm1_Gp1) * w.p;
        return (p - w.p) * dsqrt(A / (Bc + p));
    }
    return T(ph._2_Gm1) * cs(w, ph) * (pow(p / w.p, T(ph.Gm1_2G)) - 1);
}
template <class T> __device__ __forceinline__ T ex_df(T p, const P<T>& w, const Phys& ph) {
    if (p > w.p) {
        T A = T(ph._2_Gp1) / w.rho, Bc = T(ph.Gm1_Gp1) * w.p;
        return dsqrt(A / (p + Bc)) * (1 - (p - w.p) / (2 * (Bc + p)));
    }
    return pow(p / w.p, -T(ph.Gp1_2G)) / (w.rho * cs(w, ph));
}
template <class T> __device__ __forceinline__ T ex_rho(const P<T>& w, T p, const Phys& ph) {
    if (p > w.p) return w.rho * (p + T(ph.Gm1_Gp1) * w.p) / (T(ph.Gm1_Gp1) * p + w.p);
    return w.rho * pow(p / w.p, T(ph.Ginv));
}
template <class T> __device__ U<T> exact(const P<T>& L, const P<T>& R, const Phys& ph) {
    P<T> sL = pzero<T>(), sR = pzero<T>();
    sL.vy = L.vy; sL.vz = L.vz; sR.vy = R.vy; sR.vz = R.vz;
    sL.vx = L.vx; sR.vx = R.vx;
    bool done = false;
    if (ph.exact_rare_check) {
        T pmin = dmin(L.p, R.p);
        if (ex_f(pmin, L, ph) + ex_f(pmin, R, ph) + R.vx - L.vx >= T(0)) {  // two rarefactions (Exact.cpp:97-114)
            T aL = cs(L, ph), aR = cs(R, ph);
            T LR = pow(L.p / R.p, T(ph.Gm1_2G)) * aR / aL;
            sL.vx = (LR * L.vx + R.vx + T(ph._2_Gm1) * (LR * aL - aR)) / (LR + 1);
            sR.vx = sL.vx;
            sL.p = L.p * pow(1 - T(ph.Gm1_2) * (sL.vx - L.vx) / aL, T(ph._2G_Gm1));
            sR.p = sL.p;
            sL.rho = L.rho * pow(sL.p / L.p, T(ph.Ginv));
            sR.rho = R.rho * pow(sR.p / R.p, T(ph.Ginv));
            done = true;
        }
    }
    if (!done) {  // Newton iteration (Exact.cpp:74-93)
        T ps = (L.p + R.p) / T(2.0), CHA = 1;
        int it = ph.exact_iters;
        do {
            T fp = ex_f(ps, L, ph) + ex_f(ps, R, ph) + R.vx - L.vx;
            T df = ex_df(ps, L, ph) + ex_df(ps, R, ph);
            T delta = dmin(fp / df, T(0.8) * ps);
            ps -= delta;
            CHA = dabs(delta / (ps + delta / 2));
        } while (CHA > T(ph.exact_tol) && --it != 0);
        sL.p = ps; sR.p = ps;
        sL.vx = (L.vx + R.vx + ex_f(ps, R, ph) - ex_f(ps, L, ph)) / T(2.0);
        sR.vx = sL.vx;
        sL.rho = ex_rho(L, ps, ph);
        sR.rho = ex_rho(R, ps, ph);
    }
    // sample at x/t = 0 (Exact.cpp:126-167)
    P<T> st;
    T x_t = 0;
    if (dabs(sR.vx - x_t) < T(1e-12)) {
        T sql = dsqrt(sL.rho), sqr = dsqrt(sR.rho);
        U<T> m = udiv<false>(uadd<false>(umul<false>(cons<false>(sL, ph), sql), umul<false>(cons<false>(sR, ph), sqr)), sql + sqr);
        st = prim<false>(m, ph);
    } else {
        const bool left = x_t < sR.vx;
        P<T> wR = R, s = sR;
        if (left) {  // mirror
            wR = L; s = sL;
            wR.vx *= -1; s.vx *= -1;
            x_t = -x_t;
        }
        T a = cs(wR, ph);
        int zone;
        if (s.p > wR.p) {
            T scale = dsqrt(T(ph.Gp1_2G) * s.p / wR.p + T(ph.Gm1_2G));
            zone = ((x_t - wR.vx) > scale * a) ? 1 : 3;
        } else {
            if ((x_t - wR.vx) > a) zone = 1;
            else if ((x_t - s.vx) > a) zone = 2;
            else if ((x_t - s.vx) > a * pow(s.p / wR.p, T(ph.Gm1_2G))) zone = 2;
            else zone = 3;
        }
        if (zone == 1) st = wR;
        else if (zone == 3) st = s;
        else {
            st = wR;
            T scale = T(ph._2_Gp1) - T(ph.Gm1_Gp1) * (wR.vx - x_t) / a;
            st.rho *= pow(scale, T(ph._2_Gm1));
            st.p *= pow(scale, T(ph._2G_Gm1));
            st.vx = T(ph._2_Gp1) * (x_t - a + wR.vx * T(ph.Gm1_2));
        }
        if (left) st.vx *= -1;
    }
    return flux<false>(cons<false>(st, ph), st.vx, st.vy, st.vz, ph);
}

// ---- dispatch + verify-and-fallback (Flux.cpp:17-97).  Sets *fail to an
// RF_* code when no solver yields a physical update.  The runtime choice of
// Flux.cpp:34-53: MHD takes HLL, HLLE, else HLLD; hydro takes HLLC for HLLX
// and HLLC, HLL, HLLE, else the exact solver (HLLD included: it is MHD only).
template <bool MHD, class T> __device__ __forceinline__ U<T> solve(const P<T>& L, const P<T>& R, const Phys& ph) {
    if constexpr (MHD) {
        if (ph.solver == RS_HLL) return hll<true>(L, R, ph);
        if (ph.solver == RS_HLLE) return hlle<true>(L, R, ph);
        return hlld(L, R, ph);
    } else {
        if (ph.solver == RS_HLLX || ph.solver == RS_HLLC) return hllc(L, R, ph);
        if (ph.solver == RS_HLL) return hll<false>(L, R, ph);
        if (ph.solver == RS_HLLE) return hlle<false>(L, R, ph);
        return exact(L, R, ph);
    }
}
template <bool MHD, class T> __device__ __forceinline__ bool verify_ok(const P<T>& L, const P<T>& R, const U<T>& F, T dt, const Phys& ph) {
    U<T> Fd = umul<MHD>(F, dt);
    return physical<MHD>(usub<MHD>(cons<MHD>(L, ph), Fd), ph) && physical<MHD>(uadd<MHD>(cons<MHD>(R, ph), Fd), ph);
}
template <bool MHD, class T> __device__ U<T> riemann_x(const P<T>& L, const P<T>& R, T dt_dx, const Phys& ph, int* fail) {
    U<T> F = solve<MHD>(L, R, ph);
    if (ph.verify && dt_dx > T(0)) {
        T dt = dt_dx * T(ph.fallback_param);
        if (verify_ok<MHD>(L, R, F, dt, ph)) return F;
        if (ph.try_hlle) {
            F = hlle<MHD>(L, R, ph);
            if (verify_ok<MHD>(L, R, F, dt, ph)) return F;
        }
        if constexpr (!MHD) {
            if (ph.solver != RS_EXACT) {
                F = exact(L, R, ph);
                if (verify_ok<MHD>(L, R, F, dt, ph)) return F;
            }
        }
        *fail = (!physical<MHD>(L) || !physical<MHD>(R)) ? RF_BAD_INPUT : RF_BAD_SOLUTION;
    }
    return F;
}

}  // namespace dg
