#!/usr/bin/env python3
"""Generates csrc/m2p_generated.inc: straight-line FP32 M2P for the softened
Cartesian multipoles of orders p = 0, 2, 4.

Mathematics.  With Phi(y) = (|y|^2 + eps^2)^(-1/2) and y = com - target,
the reference's M2P (cpp/fmm/multipole.cpp:263-296) sums, per component c,
the exact degree-<=p Taylor coefficients of (y + x)_c (|y + x|^2 + eps^2)^(-3/2)
in the source offset x against the raw moments M_alpha about the centre of
mass.  Since y_c (|y|^2+eps^2)^(-3/2) = -d/dy_c Phi, those coefficients are
-(1/alpha!) D^(alpha + e_c) Phi(y), so

    a_c = -G * sum_alpha N_alpha D^(alpha + e_c) Phi(y),  N_alpha = M_alpha / alpha!

(N_0 = mass, N_alpha = 0 for |alpha| = 1).  For a function of |y|^2 the
derivatives factor per axis like Hermite polynomials:

    D^beta Phi = sum_gamma  prod_i beta_i! / (gamma_i! (beta_i - 2 gamma_i)! 2^gamma_i)
                            * y^(beta - 2 gamma) * T_(|beta| - |gamma|)
    T_m = (-1)^m (2m - 1)!! s^(-m - 1/2),   s = |y|^2 + eps^2.

The emitted function computes the needed q_m = s^(-m-1/2) from one rsqrt and
the monomials y^beta, then forms each derivative tensor and contracts it at
once into the three sums.
Run: python3 gen_m2p.py > m2p_generated.inc
"""
from __future__ import annotations

import itertools
from math import factorial


def exponents(max_deg):
    out = []
    for d in range(max_deg + 1):
        for a in range(d, -1, -1):
            for b in range(d - a, -1, -1):
                out.append((a, b, d - a - b))
    return out


EXPS = exponents(4)  # the 35-coefficient moment order
INDEX = {e: k for k, e in enumerate(EXPS)}


def dfact(n):
    r = 1
    while n > 1:
        r *= n
        n -= 2
    return r


def name(e):
    return f"y{e[0]}{e[1]}{e[2]}"


def deriv_terms(beta):
    """[(coef, mono, m)] with D^beta Phi = sum coef * y^mono * s^(-m-1/2)."""
    terms = []
    ranges = [range(b // 2 + 1) for b in beta]
    for g in itertools.product(*ranges):
        coef = 1.0
        for b, gi in zip(beta, g):
            coef *= factorial(b) / (factorial(gi) * factorial(b - 2 * gi) * 2 ** gi)
        m = sum(beta) - sum(g)
        coef *= (-1) ** m * dfact(2 * m - 1)
        mono = tuple(b - 2 * gi for b, gi in zip(beta, g))
        terms.append((coef, mono, m))
    return terms


def fmt(x):
    r = repr(float(x))
    return f"F({r})"


def emit(p):
    alphas = [a for a in EXPS if sum(a) == 0 or 2 <= sum(a) <= p]
    betas = set()
    for a in alphas:
        for c in range(3):
            b = list(a)
            b[c] += 1
            betas.add(tuple(b))
    betas = sorted(betas, key=lambda e: (sum(e), [-v for v in e]))
    monos = set()
    maxm = 0
    for b in betas:
        for coef, mono, m in deriv_terms(b):
            monos.add(mono)
            maxm = max(maxm, m)
    lines = []
    w = lines.append
    w(f"// order {p}: {len(alphas)} moments, {len(betas)} derivative tensors")
    w("template <typename F>")
    w(f"GC_M2P_INLINE void m2p_order{p}(F yx, F yy, F yz, F eps2, const F* __restrict__ N, F& ax,")
    w("                               F& ay, F& az) {")
    w("    const F s = GC_FMA(yz, yz, GC_FMA(yy, yy, GC_FMA(yx, yx, eps2)));")
    w("    if (!(s > F(0))) return;  // h0 == 0 contributes nothing (multipole.cpp:281-283)")
    w("    const F r = GC_RSQ(s);")
    w("    const F r2 = r * r;")
    w("    const F q0 = r;")
    for m in range(1, maxm + 1):
        w(f"    const F q{m} = q{m - 1} * r2;")
    # monomials by increasing degree, each from its predecessor (first
    # non-zero axis lowered by one) times that axis; predecessors included
    axis = ("yx", "yy", "yz")

    def pred(mono):
        for c in range(3):
            if mono[c] > 0:
                prev = list(mono)
                prev[c] -= 1
                return tuple(prev), c
        return None, None

    closure = set()
    for mono in monos:
        while mono != (0, 0, 0) and mono not in closure:
            closure.add(mono)
            mono = pred(mono)[0]
    for mono in sorted(closure, key=lambda e: (sum(e), e)):
        prev, c = pred(mono)
        if sum(prev) == 0:
            w(f"    const F {name(mono)} = {axis[c]};")
        else:
            w(f"    const F {name(mono)} = {name(prev)} * {axis[c]};")
    # each derivative tensor is contracted as soon as it is formed, so only
    # the three sums stay live (fewer registers than forming all of them)
    alpha_set = set(alphas)
    w("    F sx = F(0), sy = F(0), sz = F(0);")
    for b in betas:
        expr = None
        for coef, mono, m in deriv_terms(b):
            factor = f"{fmt(coef)} * q{m}"
            if mono == (0, 0, 0):
                term = factor
                expr = term if expr is None else f"({expr}) + {term}"
            else:
                expr = (f"({factor}) * {name(mono)}" if expr is None else
                        f"GC_FMA({factor}, {name(mono)}, {expr})")
        d = f"D{b[0]}{b[1]}{b[2]}"
        w(f"    {{ const F {d} = {expr};")
        for c, acc in enumerate(("sx", "sy", "sz")):
            if b[c] == 0:
                continue
            a = list(b)
            a[c] -= 1
            a = tuple(a)
            if a in alpha_set:
                w(f"      {acc} = GC_FMA(N[{INDEX[a]}], {d}, {acc});")
        w("    }")
    w("    ax -= sx;")
    w("    ay -= sy;")
    w("    az -= sz;")
    w("}")
    return "\n".join(lines)


def main():
    print("// m2p_generated.inc -- generated by gen_m2p.py; do not edit.")
    print("// a -= sum_alpha N[alpha] * D^(alpha+e_c) Phi(y); multiply by G outside.")
    print("// The includer defines GC_M2P_INLINE, GC_FMA(a, b, c) and GC_RSQ(s).")
    for p in (0, 2, 4):
        print(emit(p))
        print()


if __name__ == "__main__":
    main()
