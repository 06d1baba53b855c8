// The two HLLD forms of dg_math.hpp -- the reference's branching control flow
// (hlld_branch, run by the FP64 kernels) and the wavefront-uniform stream
// (hlld_uniform, run by the FP32 kernels of the mixed form) -- must return
// the same bits for every face.  Random left/right states covering every wave
// region (supersonic either way, outer star either side, double star, zero
// normal B, unphysical inputs that take the HLLE fallback) in FP64 and FP32.
// Prints one line per (type, gamma) with the region census; exit status 1 on
// any mismatch.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <random>
#include "dg_math.hpp"

using namespace dg;

template <class T> static P<T> draw(std::mt19937_64& g, int regime) {
    std::uniform_real_distribution<double> u(0, 1), s(-1, 1);
    const double rho = 0.1 + 2 * u(g), p = regime == 3 ? 1e-6 * u(g) : 0.05 + 2 * u(g);
    const double vs = regime == 1 ? 6 : 1, bs = regime == 2 ? 0 : (regime == 4 ? 4 : 1);
    P<T> w{T(rho), T(p), T(vs * s(g)), T(s(g)), T(s(g)), T(bs * s(g)), T(bs * s(g)), T(bs * s(g))};
    if (regime == 5) w.p = T(-0.1);
    return w;
}

template <class T> static int region(P<T> l, P<T> r, const Phys& ph) {
    const T Bx = (l.Bx + r.Bx) / 2;
    l.Bx = Bx; r.Bx = Bx;
    const T cl = c_fast(l, l.Bx, ph), cr = c_fast(r, r.Bx, ph);
    const T SL = dmin(l.vx - cl, r.vx - cr), SR = dmax(l.vx + cl, r.vx + cr);
    if (SL >= 0) return 0;
    if (SR <= 0) return 1;
    const T cut = dmax(dot3(l.Bx, l.By, l.Bz, l.Bx, l.By, l.Bz) + dot3(r.Bx, r.By, r.Bz, r.Bx, r.By, r.Bz), hlld_floor<T>());
    if (Bx * Bx < T(1e-12) * cut) return 2;
    const T pTL = l.p + dot3(l.Bx, l.By, l.Bz, l.Bx, l.By, l.Bz) * T(ph.k1_8pi);
    const T pTR = r.p + dot3(r.Bx, r.By, r.Bz, r.Bx, r.By, r.Bz) * T(ph.k1_8pi);
    const T xL = l.rho * (SL - l.vx), xR = r.rho * (SR - r.vx);
    const T SM = ((xR * r.vx - xL * l.vx) - (pTR - pTL)) / (xR - xL);
    const T SsL = SM - dabs(Bx) / (T(ph.sq4pi) * dsqrt(xL / (SL - SM)));
    const T SsR = SM + dabs(Bx) / (T(ph.sq4pi) * dsqrt(xR / (SR - SM)));
    return SsL >= 0 ? 3 : (SsR <= 0 ? 4 : 5);
}

template <class T> static long check(long n, double gamma, unsigned long long seed) {
    Phys ph{};
    ph.gamma = gamma; ph.gm1 = gamma - 1; ph.k1_4pi = 1 / (4 * M_PI); ph.k1_8pi = 1 / (8 * M_PI); ph.sq4pi = std::sqrt(4 * M_PI);
    ph.hlld_safety = 1;
    std::mt19937_64 g(seed);
    long bad = 0, census[6] = {0, 0, 0, 0, 0, 0};
    for (long t = 0; t < n; t++) {
        const int regime = (int)(t % 7);
        P<T> L = draw<T>(g, regime == 6 ? 0 : regime), R = draw<T>(g, regime);
        if (regime == 6) { R = L; R.vx += T(1e-3); }
        census[region(L, R, ph)]++;
        const U<T> a = hlld_uniform<T>(L, R, ph), b = hlld_branch<T>(L, R, ph);
        const T* x = &a.rho;
        const T* y = &b.rho;
        for (int k = 0; k < 8; k++) {
            const bool same = std::memcmp(x + k, y + k, sizeof(T)) == 0 || (x[k] != x[k] && y[k] != y[k]);
            if (!same) {
                if (bad < 5) printf("mismatch: case %ld regime %d field %d\n", t, regime, k);
                bad++;
                break;
            }
        }
    }
    printf("%s gamma=%.4g: %ld faces, %ld mismatches; regions supL %ld supR %ld zeroB %ld starL %ld starR %ld dstar %ld\n",
           sizeof(T) == 8 ? "fp64" : "fp32", gamma, n, bad, census[0], census[1], census[2], census[3], census[4], census[5]);
    return bad;
}

int main(int argc, char** argv) {
    const long n = argc > 1 ? atol(argv[1]) : 400000;
    long bad = check<double>(n, 5.0 / 3, 1) + check<float>(n, 5.0 / 3, 2) + check<double>(n, 1.4, 3) + check<float>(n, 1.4, 4);
    return bad ? 1 : 0;
}
