#!/usr/bin/env python3
"""Generates csrc/upward_generated.inc: straight-line FP64 P2M and M2M for
the 35 raw Cartesian moments of total degree <= 4 about the centre of mass,
so that the moments live in registers (compile-time indices).

Mathematics (cpp/fmm/multipole.cpp:226-261).  P2M adds m * o^alpha for a
particle at offset o from the cell's COM.  M2M adds a child's moments about
its COM to the parent's about the parent's COM, shifted by
h = child.com - parent.com:

    M_parent[alpha] += sum_{beta <= alpha} prod_i C(alpha_i, beta_i) h^(alpha - beta) M_child[beta]

with M_child[0] = child mass and M_child[beta] = 0 for |beta| = 1 (moments
about the COM).  Degrees 0 and 1 of the parent are its mass and 0; they are
not produced here.  The moment order is the graded order of
multipole.cpp:17-26.
Run: python3 gen_upward.py > upward_generated.inc
"""
from __future__ import annotations

from math import comb


def exponents(max_deg):
    out = []
    for d in range(max_deg + 1):
        for a in range(d, -1, -1):
            for b in range(d - a, -1, -1):
                out.append((a, b, d - a - b))
    return out


EXPS = exponents(4)
HIGH = [k for k, e in enumerate(EXPS) if sum(e) >= 2]


def mono_name(prefix, e):
    return f"{prefix}{e[0]}{e[1]}{e[2]}"


def monomials(prefix, var, out):
    """Emits prefixABC = var.x^A var.y^B var.z^C for every exponent of degree
    1..4, each from one of degree one lower times one coordinate."""
    names = {}
    axes = [f"{var}x", f"{var}y", f"{var}z"]
    for e in EXPS:
        d = sum(e)
        if d == 0:
            continue
        if d == 1:
            names[e] = axes[e.index(1)]
            continue
        # lower: remove one from the last nonzero axis
        ax = max(i for i in range(3) if e[i] > 0)
        low = list(e)
        low[ax] -= 1
        low = tuple(low)
        nm = mono_name(prefix, e)
        out.append(f"    const double {nm} = {names[low]} * {axes[ax]};")
        names[e] = nm
    return names


def emit():
    lines = []
    lines.append("// Generated by gen_upward.py; do not edit.")
    lines.append("// P2M: M[k] += m * o^e_k for the moments of degree 2..4.")
    lines.append("GC_UP_INLINE void p2m_add(double ox, double oy, double oz, double m, double* M) {")
    body = []
    names = monomials("o", "o", body)
    lines += body
    for k in HIGH:
        lines.append(f"    M[{k}] = __builtin_fma(m, {names[EXPS[k]]}, M[{k}]);")
    lines.append("}")
    lines.append("")
    lines.append("// M2M: adds a child's moments shifted by h = child COM - parent COM.")
    lines.append("GC_UP_INLINE void m2m_add(double hx, double hy, double hz, double cm,")
    lines.append("                          const double* C, double* M) {")
    body = []
    names = monomials("h", "h", body)
    lines += body
    for k in HIGH:
        a = EXPS[k]
        terms = []
        for kb, b in enumerate(EXPS):
            if any(b[i] > a[i] for i in range(3)):
                continue
            if sum(b) == 1:
                continue
            coef = comb(a[0], b[0]) * comb(a[1], b[1]) * comb(a[2], b[2])
            shift = (a[0] - b[0], a[1] - b[1], a[2] - b[2])
            operand = "cm" if sum(b) == 0 else f"C[{kb}]"
            if sum(shift) == 0:
                factor = operand
            else:
                factor = f"{names[shift]} * {operand}"
            if coef != 1:
                factor = f"{coef}.0 * {factor}" if sum(shift) == 0 else f"({coef}.0 * {names[shift]}) * {operand}"
            terms.append(factor)
        expr = terms[0]
        for t in terms[1:]:
            expr = f"({expr}) + {t}"
        lines.append(f"    M[{k}] += {expr};")
    lines.append("}")
    return "\n".join(lines) + "\n"


if __name__ == "__main__":
    print(emit(), end="")
