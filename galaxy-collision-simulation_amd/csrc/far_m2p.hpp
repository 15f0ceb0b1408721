// far_m2p.hpp — the p <= 4 Cartesian M2P of the reference
// (cpp/fmm/multipole.cpp:263-296) evaluated through the radial structure of
// the softened kernel, in about half the operations of contracting every
// derivative tensor.
//
// With Phi(y) = f(s), s = |y|^2 + eps^2, f(s) = s^(-1/2), y = com - target
// and the normalised moments N_alpha = M_alpha / alpha! (N_0 = mass,
// N_alpha = 0 for |alpha| = 1), the reference's far-field acceleration is
//     a_c = -G sum_alpha N_alpha D^(alpha + e_c) Phi(y)
//         = -G d/dy_c sum_n P_n(D) Phi(y),   P_n(x) = sum_{|alpha|=n} N_alpha x^alpha.
// Hobson's formula for a homogeneous polynomial operator acting on a
// function of |y|^2 gives
//     P_n(D) Phi = sum_k T_(n-k)(s) R_(n,k)(y),   R_(n,k) = Lap^k P_n / (2^k k!),
//     T_m = 2^m f^(m)(s) = (-1)^m (2m-1)!! s^(-m-1/2),
// and d/dy_c [T_m(s) R(y)] = T_(m+1) y_c R + T_m dR/dy_c.  Grouping by
// m = n - k into S_m = sum_(n-k=m) R_(n,k):
//     a_c = -G [ W y_c + sum_m T_m dS_m/dy_c ],   W = sum_m T_(m+1) S_m,
// with S_0 = mass, S_1 = R_21, S_2 = R_20 + R_31 + R_42, S_3 = R_30 + R_41,
// S_4 = R_40.  far_coefficients() forms, per cell and in FP64, the
// gradients of the homogeneous parts as coefficient tables (kFarCoef
// values); far_m2p() evaluates a cell at one target, taking each part's
// value from its gradient by Euler's relation y . grad R = deg(R) R.
//
// The includer defines GC_M2P_INLINE, GC_FMA(a, b, c) and GC_RSQ(s).
#pragma once

// Coefficient table of one cell (order 4 uses all, order 2 entries 0-7,
// order 0 entry 0):
//   0  S_0 = mass                       1  S_1 = R_21
//   2-7   Hessian of R_20 (xx xy xz yy yz zz)
//   8-10  gradient of R_31 (constant)   11  R_42
//   12-29 grad R_30: x, y, z rows of quadratic coefficients (xx xy xz yy yz zz)
//   30-35 Hessian of R_41 (xx xy xz yy yz zz)
//   36-65 grad R_40: x, y, z rows of cubic coefficients
//         (xxx xxy xxz xyy xyz xzz yyy yyz yzz zzz)
//   66-67 padding (zero)
constexpr int kFarCoef = 68;

// N: the 35 normalised moments in the reference's exponent order
// (cpp/fmm/multipole.cpp:12-53: degree 0..4, x exponent descending, then y).
GC_M2P_INLINE void far_coefficients(const double* __restrict__ N, double* __restrict__ K) {
#pragma clang fp contract(off)
    // degree 2: 4..9 = 200 110 101 020 011 002
    const double n200 = N[4], n110 = N[5], n101 = N[6], n020 = N[7], n011 = N[8], n002 = N[9];
    // degree 3: 10..19 = 300 210 201 120 111 102 030 021 012 003
    const double n300 = N[10], n210 = N[11], n201 = N[12], n120 = N[13], n111 = N[14],
                 n102 = N[15], n030 = N[16], n021 = N[17], n012 = N[18], n003 = N[19];
    // degree 4: 20..34 = 400 310 301 220 211 202 130 121 112 103 040 031 022 013 004
    const double n400 = N[20], n310 = N[21], n301 = N[22], n220 = N[23], n211 = N[24],
                 n202 = N[25], n130 = N[26], n121 = N[27], n112 = N[28], n103 = N[29],
                 n040 = N[30], n031 = N[31], n022 = N[32], n013 = N[33], n004 = N[34];
    K[0] = N[0];
    K[1] = (n200 + n020) + n002;
    K[2] = 2.0 * n200; K[3] = n110; K[4] = n101;
    K[5] = 2.0 * n020; K[6] = n011; K[7] = 2.0 * n002;
    K[8] = (3.0 * n300 + n120) + n102;
    K[9] = (n210 + 3.0 * n030) + n012;
    K[10] = (n201 + n021) + 3.0 * n003;
    K[11] = 3.0 * ((n400 + n040) + n004) + ((n220 + n202) + n022);
    // grad R_30
    K[12] = 3.0 * n300; K[13] = 2.0 * n210; K[14] = 2.0 * n201;
    K[15] = n120;       K[16] = n111;       K[17] = n102;
    K[18] = n210;       K[19] = 2.0 * n120; K[20] = n111;
    K[21] = 3.0 * n030; K[22] = 2.0 * n021; K[23] = n012;
    K[24] = n201;       K[25] = n111;       K[26] = 2.0 * n102;
    K[27] = n021;       K[28] = 2.0 * n012; K[29] = 3.0 * n003;
    // R_41 = Lap P_4 / 2 and its Hessian
    const double rxx = (6.0 * n400 + n220) + n202, rxy = (3.0 * n310 + 3.0 * n130) + n112,
                 rxz = (3.0 * n301 + n121) + 3.0 * n103, ryy = (n220 + 6.0 * n040) + n022,
                 ryz = (n211 + 3.0 * n031) + 3.0 * n013, rzz = (n202 + n022) + 6.0 * n004;
    K[30] = 2.0 * rxx; K[31] = rxy; K[32] = rxz;
    K[33] = 2.0 * ryy; K[34] = ryz; K[35] = 2.0 * rzz;
    // grad R_40
    K[36] = 4.0 * n400; K[37] = 3.0 * n310; K[38] = 3.0 * n301; K[39] = 2.0 * n220;
    K[40] = 2.0 * n211; K[41] = 2.0 * n202; K[42] = n130;       K[43] = n121;
    K[44] = n112;       K[45] = n103;
    K[46] = n310;       K[47] = 2.0 * n220; K[48] = n211;       K[49] = 3.0 * n130;
    K[50] = 2.0 * n121; K[51] = n112;       K[52] = 4.0 * n040; K[53] = 3.0 * n031;
    K[54] = 2.0 * n022; K[55] = n013;
    K[56] = n301;       K[57] = n211;       K[58] = 2.0 * n202; K[59] = n121;
    K[60] = 2.0 * n112; K[61] = 3.0 * n103; K[62] = n031;       K[63] = 2.0 * n022;
    K[64] = 3.0 * n013; K[65] = 4.0 * n004;
    K[66] = 0.0;
    K[67] = 0.0;
}

// a -= W y + sum_m T_m grad S_m for one cell (multiply by G outside);
// h0 = s == 0 contributes nothing (multipole.cpp:281-283): r = 0 zeroes
// every T_m, without a branch.
template <typename F, int ORDER>
GC_M2P_INLINE void far_m2p(F yx, F yy, F yz, F eps2, const F* __restrict__ K, F& ax, F& ay,
                           F& az) {
    const F s = GC_FMA(yz, yz, GC_FMA(yy, yy, GC_FMA(yx, yx, eps2)));
    const F r = s > F(0) ? GC_RSQ(s) : F(0);
    const F r2 = r * r;
    const F q1 = r * r2;
    const F t1 = -q1;
    if (ORDER == 0) {
        const F w = t1 * K[0];
        ax = GC_FMA(-w, yx, ax);
        ay = GC_FMA(-w, yy, ay);
        az = GC_FMA(-w, yz, az);
        return;
    }
    const F q2 = q1 * r2, q3 = q2 * r2;
    const F t2 = F(3) * q2, t3 = F(-15) * q3;
    // R_20: g20 = H20 y, value y.g20 / 2
    const F g20x = GC_FMA(K[4], yz, GC_FMA(K[3], yy, K[2] * yx));
    const F g20y = GC_FMA(K[6], yz, GC_FMA(K[5], yy, K[3] * yx));
    const F g20z = GC_FMA(K[7], yz, GC_FMA(K[6], yy, K[4] * yx));
    const F r20 = F(0.5) * GC_FMA(yz, g20z, GC_FMA(yy, g20y, yx * g20x));
    if (ORDER == 2) {
        const F w = GC_FMA(t3, r20, GC_FMA(t2, K[1], t1 * K[0]));
        ax -= GC_FMA(w, yx, t2 * g20x);
        ay -= GC_FMA(w, yy, t2 * g20y);
        az -= GC_FMA(w, yz, t2 * g20z);
        return;
    }
    const F q4 = q3 * r2, q5 = q4 * r2;
    const F t4 = F(105) * q4, t5 = F(-945) * q5;
    const F xx = yx * yx, xy = yx * yy, xz = yx * yz, yy2 = yy * yy, yz2 = yy * yz, zz = yz * yz;
    // S_2 = R_20 + R_31 + R_42
    const F s2 = (r20 + GC_FMA(K[10], yz, GC_FMA(K[9], yy, K[8] * yx))) + K[11];
    const F d2x = g20x + K[8], d2y = g20y + K[9], d2z = g20z + K[10];
    // S_3 = R_30 + R_41
    const F g30x = GC_FMA(K[17], zz, GC_FMA(K[16], yz2, GC_FMA(K[15], yy2,
                   GC_FMA(K[14], xz, GC_FMA(K[13], xy, K[12] * xx)))));
    const F g30y = GC_FMA(K[23], zz, GC_FMA(K[22], yz2, GC_FMA(K[21], yy2,
                   GC_FMA(K[20], xz, GC_FMA(K[19], xy, K[18] * xx)))));
    const F g30z = GC_FMA(K[29], zz, GC_FMA(K[28], yz2, GC_FMA(K[27], yy2,
                   GC_FMA(K[26], xz, GC_FMA(K[25], xy, K[24] * xx)))));
    const F g41x = GC_FMA(K[32], yz, GC_FMA(K[31], yy, K[30] * yx));
    const F g41y = GC_FMA(K[34], yz, GC_FMA(K[33], yy, K[31] * yx));
    const F g41z = GC_FMA(K[35], yz, GC_FMA(K[34], yy, K[32] * yx));
    const F d3x = g30x + g41x, d3y = g30y + g41y, d3z = g30z + g41z;
    // R_30 + R_41 = y.g30 / 3 + y.g41 / 2 = y.d3 / 3 + y.g41 / 6
    const F s3 = GC_FMA(F(1.0 / 3.0), GC_FMA(yz, d3z, GC_FMA(yy, d3y, yx * d3x)),
                        F(1.0 / 6.0) * GC_FMA(yz, g41z, GC_FMA(yy, g41y, yx * g41x)));
    // S_4 = R_40 = y.g40 / 4
    const F xxx = xx * yx, xxy = xx * yy, xxz = xx * yz, xyy = xy * yy, xyz = xy * yz;
    const F xzz = xz * yz, yyy = yy2 * yy, yyz = yy2 * yz, yzz = yz2 * yz, zzz = zz * yz;
    const F d4x = GC_FMA(K[45], zzz, GC_FMA(K[44], yzz, GC_FMA(K[43], yyz, GC_FMA(K[42], yyy,
                  GC_FMA(K[41], xzz, GC_FMA(K[40], xyz, GC_FMA(K[39], xyy, GC_FMA(K[38], xxz,
                  GC_FMA(K[37], xxy, K[36] * xxx)))))))));
    const F d4y = GC_FMA(K[55], zzz, GC_FMA(K[54], yzz, GC_FMA(K[53], yyz, GC_FMA(K[52], yyy,
                  GC_FMA(K[51], xzz, GC_FMA(K[50], xyz, GC_FMA(K[49], xyy, GC_FMA(K[48], xxz,
                  GC_FMA(K[47], xxy, K[46] * xxx)))))))));
    const F d4z = GC_FMA(K[65], zzz, GC_FMA(K[64], yzz, GC_FMA(K[63], yyz, GC_FMA(K[62], yyy,
                  GC_FMA(K[61], xzz, GC_FMA(K[60], xyz, GC_FMA(K[59], xyy, GC_FMA(K[58], xxz,
                  GC_FMA(K[57], xxy, K[56] * xxx)))))))));
    const F s4 = F(0.25) * GC_FMA(yz, d4z, GC_FMA(yy, d4y, yx * d4x));
    const F w = GC_FMA(t5, s4, GC_FMA(t4, s3, GC_FMA(t3, s2, GC_FMA(t2, K[1], t1 * K[0]))));
    ax -= GC_FMA(w, yx, GC_FMA(t4, d4x, GC_FMA(t3, d3x, t2 * d2x)));
    ay -= GC_FMA(w, yy, GC_FMA(t4, d4y, GC_FMA(t3, d3y, t2 * d2y)));
    az -= GC_FMA(w, yz, GC_FMA(t4, d4z, GC_FMA(t3, d3z, t2 * d2z)));
}
