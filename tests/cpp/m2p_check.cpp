// m2p_check.cpp — host check of the device M2P (csrc/far_m2p.hpp: per-cell
// coefficient tables and the radial-kernel evaluation) against the oracle's
// restatement of the reference M2P, in FP64 (the formula) and FP32 (the
// rounding of the device arithmetic), for well separated targets and for
// targets at about the opening-angle limit.
// Prints: max relative difference for p = 0, 2, 4 in double and float.
#include <cmath>
#include <cstdio>
#include <random>

#include "gcoracle.h"

#define GC_M2P_INLINE inline
#define GC_FMA(a, b, c) std::fma((a), (b), (c))
#define GC_RSQ(s) (F(1) / std::sqrt(s))
#include "far_m2p.hpp"

template <typename F>
void eval(int p, const double* t, const double* com, double mass, const double* mom, double eps,
          double* out) {
    int e[35][3];
    orc_multipole_exponents(&e[0][0]);
    double N[35];
    for (int k = 0; k < 35; ++k) {
        const double fact = std::tgamma(e[k][0] + 1.0) * std::tgamma(e[k][1] + 1.0) *
                            std::tgamma(e[k][2] + 1.0);
        const int d = e[k][0] + e[k][1] + e[k][2];
        N[k] = d == 0 ? mass : (d == 1 ? 0.0 : mom[k] / fact);
    }
    double K64[kFarCoef];
    far_coefficients(N, K64);
    F K[kFarCoef];
    for (int k = 0; k < kFarCoef; ++k) K[k] = F(K64[k]);
    F ax = 0, ay = 0, az = 0;
    const F y[3] = {F(com[0] - t[0]), F(com[1] - t[1]), F(com[2] - t[2])};
    const F e2 = F(eps * eps);
    if (p == 0) far_m2p<F, 0>(y[0], y[1], y[2], e2, K, ax, ay, az);
    if (p == 2) far_m2p<F, 2>(y[0], y[1], y[2], e2, K, ax, ay, az);
    if (p == 4) far_m2p<F, 4>(y[0], y[1], y[2], e2, K, ax, ay, az);
    out[0] = ax;
    out[1] = ay;
    out[2] = az;
}

int main() {
    std::mt19937_64 rng(5);
    std::normal_distribution<double> nd(0.0, 1.0);
    double worst_d[5] = {0}, worst_f[5] = {0};
    for (int trial = 0; trial < 2000; ++trial) {
        double mom[35] = {0}, com[3], t[3];
        const double size = 0.3;
        double pts[20][3], ms[20], c[3] = {0, 0, 0}, M = 0;
        for (int k = 0; k < 20; ++k) {
            ms[k] = 0.5 + 0.05 * k;
            for (int d = 0; d < 3; ++d) pts[k][d] = size * nd(rng);
            for (int d = 0; d < 3; ++d) c[d] += ms[k] * pts[k][d];
            M += ms[k];
        }
        for (int d = 0; d < 3; ++d) com[d] = c[d] / M + 2.0 * nd(rng);
        for (int k = 0; k < 20; ++k) {
            double off[3];
            for (int d = 0; d < 3; ++d) off[d] = pts[k][d] - c[d] / M;
            orc_p2m(mom, off, ms[k]);
        }
        // every other trial a target at about twice the cell's spread
        const double reach = trial % 2 ? 3.0 : 0.9;
        for (int d = 0; d < 3; ++d) t[d] = com[d] + reach * nd(rng);
        const double eps = trial % 3 == 0 ? 0.0 : 0.05;
        for (int p = 0; p <= 4; p += 2) {
            double want[3], gd[3], gf[3];
            orc_m2p(t, com, M, mom, 1.0, eps, p, want);
            eval<double>(p, t, com, M, mom, eps, gd);
            eval<float>(p, t, com, M, mom, eps, gf);
            const double nw = std::sqrt(want[0] * want[0] + want[1] * want[1] + want[2] * want[2]);
            double dd = 0, df = 0;
            for (int d = 0; d < 3; ++d) {
                dd += (gd[d] - want[d]) * (gd[d] - want[d]);
                df += (gf[d] - want[d]) * (gf[d] - want[d]);
            }
            worst_d[p] = std::fmax(worst_d[p], std::sqrt(dd) / nw);
            worst_f[p] = std::fmax(worst_f[p], std::sqrt(df) / nw);
        }
    }
    for (int p = 0; p <= 4; p += 2) std::printf("p=%d double %.3e float %.3e\n", p, worst_d[p], worst_f[p]);
    return 0;
}
