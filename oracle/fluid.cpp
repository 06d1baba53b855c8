// fluid.cpp -- oracle: state algebra, limiters, MUSCL-Hancock, Riemann solvers.
// TEST INFRASTRUCTURE ONLY (see oracle.hpp).
//
// Operation order follows the reference so that results are bit-identical
// to it when both are built without FMA contraction.  Notation "F.cpp:N"
// cites reference files relative to DRAGON/.
#include "oracle.hpp"

#include <algorithm>
#include <cmath>
#include <utility>

namespace oracle {

OConfig CFG;

// Constants.h:28-40
static const double k1_4pi = 0.25 / M_PI;
static const double k1_8pi = 0.125 / M_PI;
static const double ksq4pi = std::sqrt(4.0 * M_PI);

static inline double dot(const V3& a, const V3& b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline V3 add(V3 a, const V3& b) { a.x += b.x; a.y += b.y; a.z += b.z; return a; }
static inline V3 sub(V3 a, const V3& b) { a.x -= b.x; a.y -= b.y; a.z -= b.z; return a; }
static inline V3 mul(V3 a, double s) { a.x *= s; a.y *= s; a.z *= s; return a; }
static inline V3 divv(V3 a, double s) { a.x /= s; a.y /= s; a.z /= s; return a; }
static inline bool finite3(const V3& a) { return std::isfinite(a.x) && std::isfinite(a.y) && std::isfinite(a.z); }

// Cons arithmetic (Arithmetic.cpp:14-128): component-wise, rho, mom, E, B.
static inline Cons cadd(Cons a, const Cons& b) {
    a.rho += b.rho; a.mom = add(a.mom, b.mom); a.E += b.E;
    if (CFG.mhd) a.B = add(a.B, b.B);
    return a;
}
static inline Cons csub(Cons a, const Cons& b) {
    a.rho -= b.rho; a.mom = sub(a.mom, b.mom); a.E -= b.E;
    if (CFG.mhd) a.B = sub(a.B, b.B);
    return a;
}
static inline Cons cmul(Cons a, double s) {
    a.rho *= s; a.mom = mul(a.mom, s); a.E *= s;
    if (CFG.mhd) a.B = mul(a.B, s);
    return a;
}
static inline Cons cdiv(Cons a, double s) {
    a.rho /= s; a.mom = divv(a.mom, s); a.E /= s;
    if (CFG.mhd) a.B = divv(a.B, s);
    return a;
}

// ------------------------------------------------------ FluidElement.cpp
double energy(const Prim& w) {  // FluidElement.cpp:53-59
    double E = w.p / (CFG.gamma - 1.0) + (w.rho / 2) * dot(w.v, w.v);
    if (CFG.mhd) E += dot(w.B, w.B) * k1_8pi;
    return E;
}
double pressure(const Cons& u) {  // FluidElement.cpp:60-66
    double Et = u.E - dot(u.mom, u.mom) / (2 * u.rho);
    if (CFG.mhd) Et -= dot(u.B, u.B) * k1_8pi;
    return (CFG.gamma - 1.0) * Et;
}
Cons to_cons(const Prim& w) {  // FluidElement.cpp:35-42
    Cons u;
    u.rho = w.rho;
    u.mom = mul(w.v, w.rho);
    u.E = energy(w);
    if (CFG.mhd) u.B = w.B;
    return u;
}
Prim to_prim(const Cons& u) {  // FluidElement.cpp:44-51
    Prim w;
    w.rho = u.rho;
    w.v = divv(u.mom, u.rho);
    w.p = pressure(u);
    if (CFG.mhd) w.B = u.B;
    return w;
}
double enthalpy(const Prim& w) {  // FluidElement.cpp:69-75
    double h = energy(w) + w.p;
    if (CFG.mhd) h += dot(w.B, w.B) * k1_8pi;
    return h / w.rho;
}
bool is_physical(const Prim& w) {  // FluidElement.cpp:82-87
    if (CFG.mhd && !finite3(w.B)) return false;
    return std::isfinite(w.rho) && finite3(w.v) && std::isfinite(w.p) && w.rho > 0.0 && w.p > 0.0;
}
bool is_finite(const Cons& u) {  // FluidElement.cpp:88-93
    if (CFG.mhd && !finite3(u.B)) return false;
    return std::isfinite(u.rho) && finite3(u.mom) && std::isfinite(u.E);
}
bool is_physical(const Cons& u) { return is_finite(u) && u.rho > 0.0 && pressure(u) > 0.0; }  // :94-96

double cs(const Prim& w) { return std::sqrt(CFG.gamma * w.p / w.rho); }  // :99-101
double c_alfven(const Prim& w) { return std::sqrt(dot(w.B, w.B) / w.rho) / ksq4pi; }  // :103-105
double c_fast(const Prim& w, double Bk) {  // :107-114 (pow(x,2) is x*x)
    double c = cs(w), a = c_alfven(w), c2 = c * c, a2 = a * a, ak2 = (Bk * Bk / w.rho) * k1_4pi;
    double s = c2 + a2;
    double disc = std::max(0.0, (s * s) * 0.25 - ak2 * c2);
    return std::sqrt((c2 + a2) * 0.5 + std::sqrt(disc));
}
double c_fast_x(const Prim& w) { return c_fast(w, w.B.x); }
double c_fast_max(const Prim& w) {  // :115-120
    double Bk;
    if (w.B.x * w.B.x < w.B.y * w.B.y) Bk = (w.B.x * w.B.x < w.B.z * w.B.z) ? w.B.x : w.B.z;
    else Bk = (w.B.y * w.B.y < w.B.z * w.B.z) ? w.B.y : w.B.z;
    return c_fast(w, Bk);
}

Cons flux_v(const Cons& u, const V3& v) {  // FluidElement.cpp:133-155
    Cons F;
    double p = pressure(u);
    F.rho = u.mom.x;
    F.mom = mul(u.mom, v.x);
    F.mom.x += p;
    F.E = (u.E + p) * v.x;
    if (CFG.mhd) {
        double pmag = dot(u.B, u.B) * k1_8pi;
        F.mom.x += pmag;
        F.E += pmag * v.x;
        V3 ten = mul(mul(u.B, u.B.x), k1_4pi);
        F.mom = sub(F.mom, ten);
        F.E -= dot(v, ten);
        F.B = sub(mul(u.B, v.x), mul(v, u.B.x));
        F.B.x = 0;
    }
    return F;
}
Cons flux_cons(const Cons& u) {  // :125-132
    if (CFG.mhd) return flux_v(u, divv(u.mom, u.rho));
    V3 v; v.x = u.mom.x / u.rho;
    return flux_v(u, v);
}
Cons flux_prim(const Prim& w) { return flux_v(to_cons(w), w.v); }  // :124

Prim swapXY(Prim w) { std::swap(w.v.x, w.v.y); if (CFG.mhd) std::swap(w.B.x, w.B.y); return w; }
Prim swapXZ(Prim w) { std::swap(w.v.x, w.v.z); if (CFG.mhd) std::swap(w.B.x, w.B.z); return w; }
Cons swapXY(Cons u) { std::swap(u.mom.x, u.mom.y); if (CFG.mhd) std::swap(u.B.x, u.B.y); return u; }
Cons swapXZ(Cons u) { std::swap(u.mom.x, u.mom.z); if (CFG.mhd) std::swap(u.B.x, u.B.z); return u; }

// ---------------------------------------------------------- Limiters.cpp
double minmod(double a, double b) {  // :88-91
    if (a * b <= 0) return 0;
    return (std::abs(a) < std::abs(b)) ? a : b;
}
double lim_mc(double a, double b) {  // :112-117
    if (a * b <= 0) return 0;
    double c = std::abs(a + b);
    if (c < 4 * std::abs(a) && c < 4 * std::abs(b)) return a < 0 ? -c / 2 : c / 2;
    return 2 * minmod(a, b);
}
double lim_vanleer(double a, double b) {  // :138-141
    if (a * b <= 0) return 0;
    return 2 * a * b / (a + b);
}
double lim_superbee(double a, double b) {  // :162-168
    if (a * b <= 0) return 0;
    double A = std::abs(a), Bb = std::abs(b);
    double s1 = std::min(2 * A, Bb), s2 = std::min(A, 2 * Bb);
    double m = std::max(s1, s2);
    return a < 0 ? -m : m;
}
double lim_vanalbada(double a, double b) {  // :188-191
    if (a * b <= 0) return 0;
    return a * b * (a + b) / (a * a + b * b);
}
double limit1(double a, double b) {  // :17-42 dispatch (unknown => minmod)
    switch (CFG.limiter) {
        case LIM_MC: return lim_mc(a, b);
        case LIM_VANLEER: return lim_vanleer(a, b);
        case LIM_SUPERBEE: return lim_superbee(a, b);
        case LIM_VANALBADA: return lim_vanalbada(a, b);
        default: return minmod(a, b);
    }
}
Prim limit(const Prim& a, const Prim& b) {
    Prim W;
    W.rho = limit1(a.rho, b.rho);
    W.v.x = limit1(a.v.x, b.v.x); W.v.y = limit1(a.v.y, b.v.y); W.v.z = limit1(a.v.z, b.v.z);
    W.p = limit1(a.p, b.p);
    if (CFG.mhd) { W.B.x = limit1(a.B.x, b.B.x); W.B.y = limit1(a.B.y, b.B.y); W.B.z = limit1(a.B.z, b.B.z); }
    return W;
}

// ------------------------------------------------------------- MUSCL.cpp
void muscl(const Prim& wL, Prim& L, const Prim& wC, Prim& R, const Prim& wR, double dt_dL, const V3& dB) {
    if (CFG.muscl) {
        Prim ld, rd;
        ld.rho = wC.rho - wL.rho; rd.rho = wR.rho - wC.rho;
        ld.v = sub(wC.v, wL.v); rd.v = sub(wR.v, wC.v);
        ld.p = wC.p - wL.p; rd.p = wR.p - wC.p;
        if (CFG.mhd) { ld.B = sub(wC.B, wL.B); rd.B = sub(wR.B, wC.B); }
        Prim dW = limit(ld, rd);
        Prim l, r;
        l.rho = wC.rho - 0.5 * dW.rho;
        r.rho = wC.rho + 0.5 * dW.rho;
        l.v = sub(wC.v, mul(dW.v, 0.5));
        r.v = add(wC.v, mul(dW.v, 0.5));
        l.p = wC.p - 0.5 * dW.p;
        r.p = wC.p + 0.5 * dW.p;
        if (CFG.mhd) { l.B = sub(wC.B, mul(dW.B, 0.5)); r.B = add(wC.B, mul(dW.B, 0.5)); }
        Cons UL = to_cons(l), UR = to_cons(r);
        Cons corr = cmul(csub(flux_v(UR, r.v), flux_v(UL, l.v)), 0.5 * dt_dL);
        if (CFG.mhd) {  // MUSCL.cpp:255-264
            const double a_s = (r.B.x - l.B.x) * dt_dL;
            const double vy = 0.5 * (l.v.y + r.v.y), vz = 0.5 * (l.v.z + r.v.z);
            const double dBy_ = 0.5 * wC.v.y * minmod(dB.x, -dB.y) - 0.5 * vy * a_s;
            const double dBz_ = 0.5 * wC.v.z * minmod(dB.x, -dB.z) - 0.5 * vz * a_s;
            corr.B.y -= dBy_;
            corr.B.z -= dBz_;
            corr.E -= k1_4pi * (wC.B.y * dBy_ + wC.B.z * dBz_);
        }
        l = to_prim(csub(UL, corr));
        r = to_prim(csub(UR, corr));
        if (is_physical(l) && is_physical(r)) { L = l; R = r; return; }
    }
    L = wC;
    R = wC;
}

// --------------------------------------------------------------- HLL.cpp
Prim roe_avg(const Prim& L, const Prim& R) {  // HLL.cpp:21-37
    double sql = std::sqrt(L.rho), sqr = std::sqrt(R.rho);
    double _sql = sql / (sql + sqr), _sqr = sqr / (sql + sqr);
    Prim w;
    w.rho = sql * sqr;
    w.v = add(mul(L.v, _sql), mul(R.v, _sqr));
    if (CFG.mhd) w.B = add(mul(L.B, _sql), mul(R.B, _sqr));
    double H = _sql * enthalpy(L) + _sqr * enthalpy(R);
    double a2 = (CFG.gamma - 1) * (H - dot(w.v, w.v) / 2);
    w.p = a2 * w.rho / CFG.gamma;
    return w;
}

Cons Riemann::HLL() {  // HLL.cpp:42-60
    Prim M = roe_avg(L, R);
    double aL, aR, aM;
    if (CFG.mhd) {
        double Bx = (L.B.x + R.B.x) / 2;
        L.B.x = Bx; R.B.x = Bx;
        aL = c_fast_x(L); aR = c_fast_x(R); aM = c_fast_x(M);
    } else {
        aL = cs(L); aR = cs(R); aM = cs(M);
    }
    double SL = std::min(L.v.x - aL, M.v.x - aM);
    double SR = std::max(R.v.x + aR, M.v.x + aM);
    return HLL(SL, SR);
}
Cons Riemann::HLLE() {  // HLL.cpp:61-83
    double sql = std::sqrt(L.rho), sqr = std::sqrt(R.rho);
    double _sql = sql / (sql + sqr), _sqr = sqr / (sql + sqr);
    double aL, aR;
    if (CFG.mhd) {
        double Bx = (L.B.x + R.B.x) / 2;
        L.B.x = Bx; R.B.x = Bx;
        aL = c_fast_x(L); aR = c_fast_x(R);
    } else {
        aL = cs(L); aR = cs(R);
    }
    double _v = (R.v.x - L.v.x) / (sql + sqr);
    double d = (_sql * aL * aL + _sqr * aR * aR) + 0.5 * (sql * sqr) * _v * _v;
    d = std::sqrt(d);
    double u = _sql * L.v.x + _sqr * R.v.x;
    double SL = std::min(L.v.x - aL, u - d);
    double SR = std::max(R.v.x + aR, u + d);
    return HLL(SL, SR);
}
Cons Riemann::HLL(double sl, double sr) {  // HLL.cpp:84-96
    if (sl >= 0) return flux_v(to_cons(L), L.v);
    if (sr <= 0) return flux_v(to_cons(R), R.v);
    Cons UL = to_cons(L), UR = to_cons(R);
    Cons FL = flux_v(UL, L.v), FR = flux_v(UR, R.v);
    return cdiv(cadd(csub(cmul(FL, sr), cmul(FR, sl)), cmul(csub(UR, UL), sl * sr)), sr - sl);
}
Cons Riemann::HLLC() {  // HLL.cpp:103-112
    Prim M = roe_avg(L, R);
    double aL = cs(L), aR = cs(R), aM = cs(M);
    double SL = std::min(L.v.x - aL, M.v.x - aM);
    double SR = std::max(R.v.x + aR, M.v.x + aM);
    return HLLC(SL, SR);
}
Cons Riemann::HLLC(double sl, double sr) {  // HLL.cpp:114-132
    if (sl >= 0) return flux_prim(L);
    if (sr <= 0) return flux_prim(R);
    double _pr = R.p + R.rho * R.v.x * (R.v.x - sr), _pl = L.p + L.rho * L.v.x * (L.v.x - sl);
    double sc = (_pr - _pl) / (R.rho * (R.v.x - sr) - L.rho * (L.v.x - sl));
    const Prim& X = (sc > 0 ? L : R);
    double sx = sc > 0 ? sl : sr;
    Cons UX = to_cons(X);
    Cons U = cdiv(cmul(UX, sx - X.v.x), sx - sc);
    U.mom.x = U.rho * sc;
    U.E += U.rho * (sc - X.v.x) * (sc + X.p / (X.rho * (sx - X.v.x)));
    return cadd(flux_v(UX, X.v), cmul(csub(U, UX), sx));
}

// -------------------------------------------------------------- HLLD.cpp
static void outer_star(Cons& us, const Prim& K, double xK, double SK, double SM, double pT, double Bx) {
    double Bx2 = (Bx * Bx) * k1_4pi;  // HLLD.cpp:21-37
    double den = xK * (SK - SM) - Bx2;
    bool tiny = std::abs(den) < 1e-12;
    us.B = tiny ? K.B : divv(mul(K.B, xK * (SK - K.v.x) - Bx2), den);
    us.B.x = Bx;
    // K.B * Bx * (SM - vx) / denom * 1/4pi, evaluated left to right
    V3 vs = tiny ? sub(K.v, V3{}) : sub(K.v, mul(divv(mul(mul(K.B, Bx), SM - K.v.x), den), k1_4pi));
    vs.x = SM;
    us.mom = mul(vs, us.rho);
    us.E = energy(K) * (SK - K.v.x);
    us.E += pT * SM - (K.p + dot(K.B, K.B) * k1_8pi) * K.v.x;
    us.E += Bx * (dot(K.v, K.B) - dot(vs, us.B)) * k1_4pi;
    us.E /= (SK - SM);
}
Cons Riemann::HLLD() {  // HLLD.cpp:38-133
    double Bx = (L.B.x + R.B.x) / 2;
    L.B.x = Bx; R.B.x = Bx;
    double cl = c_fast_x(L), cr = c_fast_x(R);
    double SL = std::min(L.v.x - cl, R.v.x - cr);
    if (SL >= 0) return flux_prim(L);
    double SR = std::max(L.v.x + cl, R.v.x + cr);
    if (SR <= 0) return flux_prim(R);
    double cutoff = std::max(dot(L.B, L.B) + dot(R.B, R.B), 1e-100);
    if (Bx * Bx < 1e-12 * cutoff) return HLLD_zero_B(SL, SR);

    double pTL = L.p + dot(L.B, L.B) * k1_8pi, pTR = R.p + dot(R.B, R.B) * k1_8pi;
    double xL = L.rho * (SL - L.v.x), xR = R.rho * (SR - R.v.x);
    double SM = ((xR * R.v.x - xL * L.v.x) - (pTR - pTL)) / (xR - xL);
    Cons usL, usR;
    usL.rho = xL / (SL - SM);
    double SsL = SM - std::abs(Bx) / (ksq4pi * std::sqrt(usL.rho));
    usR.rho = xR / (SR - SM);
    double SsR = SM + std::abs(Bx) / (ksq4pi * std::sqrt(usR.rho));
    Cons FsL, FsR;
    Prim wsL, wsR;
    double pT = (xR * pTL - xL * pTR + xL * xR * (R.v.x - L.v.x)) / (xR - xL);
    if (SsR >= 0) {
        outer_star(usL, L, xL, SL, SM, pT, Bx);
        wsL.rho = usL.rho; wsL.v = divv(usL.mom, usL.rho); wsL.B = usL.B;
        if (CFG.hlld_safety) {
            wsL.p = pT - dot(wsL.B, wsL.B) * k1_8pi;
            if (!is_physical(wsL) || !is_physical(usL)) return HLLE();
        }
        FsL = cadd(flux_prim(L), cmul(csub(usL, to_cons(L)), SL)); FsL.B.x = 0;
    }
    if (SsL <= 0) {
        outer_star(usR, R, xR, SR, SM, pT, Bx);
        wsR.rho = usR.rho; wsR.v = divv(usR.mom, usR.rho); wsR.B = usR.B;
        if (CFG.hlld_safety) {
            wsR.p = pT - dot(wsR.B, wsR.B) * k1_8pi;
            if (!is_physical(wsR) || !is_physical(usR)) return HLLE();
        }
        FsR = cadd(flux_prim(R), cmul(csub(usR, to_cons(R)), SR)); FsR.B.x = 0;
    }
    if (SsL >= 0) return FsL;
    if (SsR <= 0) return FsR;
    Prim ws = SM >= 0 ? wsL : wsR;
    double sqL = std::sqrt(wsL.rho), sqR = std::sqrt(wsR.rho);
    double _sqL = sqL / (sqL + sqR), _sqR = sqR / (sqL + sqR);
    double sbx = Bx > 0 ? ksq4pi : (Bx < 0 ? -ksq4pi : 0);
    Cons uss;
    Prim wss;
    wss.rho = ws.rho;
    uss.rho = wss.rho;
    wss.v = add(add(mul(wsL.v, _sqL), mul(wsR.v, _sqR)), divv(mul(mul(sub(usR.B, usL.B), sbx), k1_4pi), sqL + sqR));
    wss.v.x = SM;
    uss.mom = mul(wss.v, uss.rho);
    wss.B = add(add(mul(usR.B, _sqL), mul(usL.B, _sqR)), divv(mul(mul(mul(sub(wsR.v, wsL.v), sbx), sqL), sqR), sqL + sqR));
    wss.B.x = Bx;
    uss.B = wss.B;
    double dE = sbx * (dot(ws.v, ws.B) - dot(wss.v, wss.B)) * k1_4pi * std::sqrt(wss.rho);
    uss.E = SM >= 0 ? usL.E - dE : usR.E + dE;
    if (CFG.hlld_safety) {
        wss.p = pT - dot(wss.B, wss.B) * k1_8pi;
        if (!is_physical(wss) || !is_finite(uss)) return HLLE();
    }
    if (SM >= 0) return cadd(FsL, cmul(csub(uss, usL), SsL));
    return cadd(FsR, cmul(csub(uss, usR), SsR));
}
static void outer_star_zero(Cons& us, const Prim& K, double SK, double SM, double pT) {  // HLLD.cpp:137-150
    us.B = (std::abs(SK - SM) < 1e-12) ? K.B : divv(mul(K.B, SK - K.v.x), SK - SM);
    us.B.x = 0;
    V3 vs = K.v;
    vs.x = SM;
    us.mom = mul(vs, us.rho);
    us.E = energy(K) * (SK - K.v.x);
    us.E += pT * SM - (K.p + dot(K.B, K.B) * k1_8pi) * K.v.x;
    us.E /= (SK - SM);
}
Cons Riemann::HLLD_zero_B(double SL, double SR) {  // HLLD.cpp:152-205
    double pTL = L.p + dot(L.B, L.B) * k1_8pi, pTR = R.p + dot(R.B, R.B) * k1_8pi;
    double xL = L.rho * (SL - L.v.x), xR = R.rho * (SR - R.v.x);
    double SM = ((xR * R.v.x - xL * L.v.x) - (pTR - pTL)) / (xR - xL);
    Cons usL, usR;
    usL.rho = xL / (SL - SM);
    usR.rho = xR / (SR - SM);
    Cons FsL, FsR;
    Prim wsL, wsR;
    double pT = (xR * pTL - xL * pTR + xL * xR * (R.v.x - L.v.x)) / (xR - xL);
    if (SM >= 0) {
        outer_star_zero(usL, L, SL, SM, pT);
        wsL.rho = usL.rho; wsL.v = divv(usL.mom, usL.rho); wsL.B = usL.B;
        if (CFG.hlld_safety) {
            wsL.p = pT - dot(wsL.B, wsL.B) * k1_8pi;
            if (!is_physical(wsL) || !is_finite(usL)) return HLLE();
        }
        FsL = cadd(flux_prim(L), cmul(csub(usL, to_cons(L)), SL)); FsL.B.x = 0;
    }
    if (SM <= 0) {
        outer_star_zero(usR, R, SR, SM, pT);
        wsR.rho = usR.rho; wsR.v = divv(usR.mom, usR.rho); wsR.B = usR.B;
        if (CFG.hlld_safety) {
            wsR.p = pT - dot(wsR.B, wsR.B) * k1_8pi;
            if (!is_physical(wsR) || !is_finite(usR)) return HLLE();
        }
        FsR = cadd(flux_prim(R), cmul(csub(usR, to_cons(R)), SR)); FsR.B.x = 0;
    }
    if (SM > 0) return FsL;
    if (SM < 0) return FsR;
    double sql = std::sqrt(wsL.rho), sqr = std::sqrt(wsR.rho);
    double _sql = sql / (sql + sqr), _sqr = sqr / (sql + sqr);
    return cadd(cmul(FsL, _sql), cmul(FsR, _sqr));
}

// ------------------------------------------------------------- Exact.cpp
static inline double G(int which) {
    const double g = CFG.gamma;
    switch (which) {
        case 0: return (g - 1.0) / (2.0 * g);   // _Gm1_2G
        case 1: return (g + 1.0) / (2.0 * g);   // _Gp1_2G
        case 2: return 2.0 * g / (g - 1.0);     // _2G_Gm1
        case 3: return (g - 1.0) / 2.0;         // _Gm1_2
        case 4: return 2.0 / (g - 1.0);         // _2_Gm1
        case 5: return 2.0 / (g + 1.0);         // _2_Gp1
        case 6: return (g - 1.0) / (g + 1.0);   // _Gm1_Gp1
        default: return 1.0 / g;                // _Ginv
    }
}
double Riemann::f(double p, const Prim& w) {  // Exact.cpp:29-37
    if (p > w.p) {
        double A = G(5) / w.rho, Bc = G(6) * w.p;
        return (p - w.p) * std::sqrt(A / (Bc + p));
    }
    return G(4) * cs(w) * (std::pow(p / w.p, G(0)) - 1);
}
static double df(double p, const Prim& w) {  // Exact.cpp:38-45
    if (p > w.p) {
        double A = G(5) / w.rho, Bc = G(6) * w.p;
        return std::sqrt(A / (p + Bc)) * (1 - (p - w.p) / (2 * (Bc + p)));
    }
    return std::pow(p / w.p, -G(1)) / (w.rho * cs(w));
}
double Riemann::exact_StarP(double pGuess) {  // Exact.cpp:74-85
    double pStar = pGuess, CHA = 1;
    int iters = CFG.exact_max_iters;
    do {
        double fp = f(pStar, L) + f(pStar, R) + R.v.x - L.v.x;
        double dfdp = df(pStar, L) + df(pStar, R);
        double delta = std::min(fp / dfdp, 0.8 * pStar);
        pStar -= delta;
        CHA = std::abs(delta / (pStar + delta / 2));
    } while (CHA > CFG.exact_tol && --iters != 0);
    return pStar;
}
double Riemann::exact_StarV(double pStar) { return (L.v.x + R.v.x + f(pStar, R) - f(pStar, L)) / 2.0; }
static double star_rho(const Prim& w, double p) {  // Exact.cpp:90-93
    if (p > w.p) return w.rho * (p + G(6) * w.p) / (G(6) * p + w.p);
    return w.rho * std::pow(p / w.p, G(7));
}
void Riemann::exact_solution(Prim& sL, Prim& sR, bool& trrs) {  // Exact.cpp:47-70, 97-114
    sL = Prim(); sR = Prim();
    sL.v = L.v; sR.v = R.v;  // RiemannSolution ctor copies transverse values
    trrs = false;
    if (CFG.exact_rarefaction_check) {
        double p_min = std::min(L.p, R.p);
        if (f(p_min, L) + f(p_min, R) + R.v.x - L.v.x >= 0) {
            trrs = true;
            double aL = cs(L), aR = cs(R);
            double LR = std::pow(L.p / R.p, G(0)) * aR / aL;
            sL.v.x = (LR * L.v.x + R.v.x + G(4) * (LR * aL - aR)) / (LR + 1);
            sR.v.x = sL.v.x;
            sL.p = L.p * std::pow(1 - G(3) * (sL.v.x - L.v.x) / aL, G(2));
            sR.p = sL.p;
            sL.rho = L.rho * std::pow(sL.p / L.p, G(7));
            sR.rho = R.rho * std::pow(sR.p / R.p, G(7));
            return;
        }
    }
    sL.p = exact_StarP((L.p + R.p) / 2.0);
    sR.p = sL.p;
    sL.v.x = exact_StarV(sL.p);
    sR.v.x = sL.v.x;
    sL.rho = star_rho(L, sL.p);
    sR.rho = star_rho(R, sR.p);
}
static void mirror(Prim& wL, Prim& wR, Prim& sL, Prim& sR) {  // Exact.cpp:169-184
    std::swap(wL, wR);
    std::swap(sL, sR);
    wL.v.x *= -1; wR.v.x *= -1; sL.v.x *= -1; sR.v.x *= -1;
    if (CFG.mhd) { wL.B.x *= -1; wR.B.x *= -1; sL.B.x *= -1; sR.B.x *= -1; }
}
Prim exact_sample(const Prim& wL0, const Prim& wR0, const Prim& sL0, const Prim& sR0, double x_t) {
    // Exact.cpp:126-167
    Prim wL = wL0, wR = wR0, sL = sL0, sR = sR0;
    if (std::abs(sR.v.x - x_t) < 1e-12) {
        double sql = std::sqrt(sL.rho), sqr = std::sqrt(sR.rho);
        return to_prim(cdiv(cadd(cmul(to_cons(sL), sql), cmul(to_cons(sR), sqr)), sql + sqr));
    }
    bool isLeft = x_t < sR.v.x;
    if (isLeft) { mirror(wL, wR, sL, sR); x_t = -x_t; }
    double a = cs(wR);
    int zone = 0;
    if (sR.p > wR.p) {
        double scale = std::sqrt(G(1) * sR.p / wR.p + G(0));
        zone = ((x_t - wR.v.x) > scale * a) ? 1 : 3;
    } else {
        if ((x_t - wR.v.x) > a) zone = 1;
        else if ((x_t - sR.v.x) > a) zone = 2;
        else if ((x_t - sR.v.x) > a * std::pow(sR.p / wR.p, G(0))) zone = 2;
        else zone = 3;
    }
    Prim st;
    switch (zone) {
        case 1: st = wR; break;
        case 3: st = sR; break;
        default: {
            st = wR;
            double scale = G(5) - G(6) * (wR.v.x - x_t) / a;
            st.rho *= std::pow(scale, G(4));
            st.p *= std::pow(scale, G(2));
            st.v.x = G(5) * (x_t - a + wR.v.x * G(3));
        }
    }
    if (isLeft) { st.v.x *= -1; }
    return st;
}
Cons Riemann::exact_flux() {  // Exact.cpp:120-123 with x/t = 0
    Prim sL, sR;
    bool trrs;
    exact_solution(sL, sR, trrs);
    Prim w = exact_sample(L, R, sL, sR, 0.0);
    return flux_v(to_cons(w), w.v);
}

// -------------------------------------------------------------- Flux.cpp
Cons Riemann::flux(double dt_dx) {  // Flux.cpp:17-74 (CHOOSE_RUNTIME semantics)
    Cons F;
    int s = CFG.solver;
    if (CFG.mhd) {
        switch (s) {
            case RS_HLL: F = HLL(); break;
            case RS_HLLE: F = HLLE(); break;
            default: F = HLLD(); break;
        }
    } else {
        switch (s) {
            case RS_HLLX: case RS_HLLC: F = HLLC(); break;
            case RS_HLL: F = HLL(); break;
            case RS_HLLE: F = HLLE(); break;
            default: F = exact_flux(); break;
        }
    }
    if (CFG.verify() && dt_dx > 0) verify_and_fallback(F, dt_dx);
    return F;
}
void Riemann::verify_and_fallback(Cons& F, double dt_dx) {  // Flux.cpp:77-97
    dt_dx *= CFG.fallback_param;
    auto ok = [&](const Cons& G) {
        return is_physical(csub(to_cons(L), cmul(G, dt_dx))) && is_physical(cadd(to_cons(R), cmul(G, dt_dx)));
    };
    if (ok(F)) return;
    if (CFG.fallback_try_hlle) {
        F = HLLE();
        if (ok(F)) return;
    }
    if (!CFG.mhd && CFG.solver != RS_EXACT) {
        F = exact_flux();
        if (ok(F)) return;
    }
    if (!is_physical(L) || !is_physical(R))
        throw std::runtime_error("Riemann Solution is unphysical because inputs were not physical");
    throw std::runtime_error("Riemann Solution to physical inputs is unphysical");
}
Cons Riemann::flux_Y(double dt_dy) {  // Flux.cpp:102-108
    Riemann r(swapXY(L), swapXY(R));
    return swapXY(r.flux(dt_dy));
}
Cons Riemann::flux_Z(double dt_dz) {  // Flux.cpp:109-115
    Riemann r(swapXZ(L), swapXZ(R));
    return swapXZ(r.flux(dt_dz));
}

}  // namespace oracle
