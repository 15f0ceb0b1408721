// m2p_check.cpp — host check of the generated M2P (csrc/m2p_generated.inc)
// against the oracle's restatement of the reference M2P, in FP64 (the
// formula) and FP32 (the rounding of the device arithmetic).
// Prints: max relative difference for p = 0, 2, 4 in double and float.
#include <cmath>
#include <cstdio>
#include <random>

#include "gcoracle.h"

#define GC_M2P_INLINE inline
#define GC_FMA(a, b, c) std::fma((a), (b), (c))
#define GC_RSQ(s) (F(1) / std::sqrt(s))
#include "m2p_generated.inc"

template <typename F>
void eval(int p, const double* t, const double* com, double mass, const double* mom, double eps,
          double* out) {
    int e[35][3];
    orc_multipole_exponents(&e[0][0]);
    F N[35];
    for (int k = 0; k < 35; ++k) {
        const double fact = std::tgamma(e[k][0] + 1.0) * std::tgamma(e[k][1] + 1.0) *
                            std::tgamma(e[k][2] + 1.0);
        const int d = e[k][0] + e[k][1] + e[k][2];
        N[k] = F(d == 0 ? mass : (d == 1 ? 0.0 : mom[k] / fact));
    }
    F ax = 0, ay = 0, az = 0;
    const F y[3] = {F(com[0] - t[0]), F(com[1] - t[1]), F(com[2] - t[2])};
    const F e2 = F(eps * eps);
    if (p == 0) m2p_order0<F>(y[0], y[1], y[2], e2, N, ax, ay, az);
    if (p == 2) m2p_order2<F>(y[0], y[1], y[2], e2, N, ax, ay, az);
    if (p == 4) m2p_order4<F>(y[0], y[1], y[2], e2, N, ax, ay, az);
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
        for (int d = 0; d < 3; ++d) t[d] = com[d] + 3.0 * nd(rng);
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
