"""Generates epr_wedge_tables.h: exact element integrals of the wedge energy.

For the multilinear hat basis on the unit element [0,1]^Dim, with corners
indexed by bits (coordinate 0 most significant, the reference's relative index
of Include/Hat.inl:462-477), the wedge mass tensor is

    M(a,b,c,d) = int <grad phi_a ^ grad phi_b , grad phi_c ^ grad phi_d>

with the reference's 2-form inner product (Include/Hat.inl:32-37, 240-248,
358-370): for skew packing p_ij = (u_j v_i - u_i v_j)/2 (i>j) the inner
product is the full Frobenius product, 2 * sum_{i>j} p_ij q_ij.  At grid
resolution r every entry scales by r^(4-Dim) (four derivatives, one volume).

M is invariant under reflecting the element (corner index XOR a mask), so the
Gauss-Seidel fit of a node only needs the tables of corner 0:

    WT[j][kl] = M(0, j, k, l)      j in [1,2^Dim), k<l   (cubic terms)
    WA[j][k]  = M(0, j, 0, k)      j,k in [1,2^Dim)      (quadratic terms)

It also emits the full-stiffness element table used to assemble B
(Include/Hat.scalars.inl:550-565) at unit resolution,

    FS[i1][i2][a][b] = int d_b phi_i1 * d_a phi_i2,

evaluated with the reference's floating-point operation order
(Include/Hat.inl:188-194, ElementFunction::integral), so that the table
scaled by r^(2-Dim) (exact: r is a power of two) equals the reference's.

Run: python gen_wedge_tables.py > epr_wedge_tables.h
"""
from fractions import Fraction
from math import factorial


def integral_1d(p, q):
    """int_0^1 x^p (1-x)^q dx = p! q! / (p+q+1)!"""
    return Fraction(factorial(p) * factorial(q), factorial(p + q + 1))


def tables(dim):
    C = 1 << dim

    def bit(c, d):
        return (c >> (dim - 1 - d)) & 1

    def grad(c, a):
        # d/dt_a of prod_d psi_{c_d}(t_d), psi_0 = 1-t, psi_1 = t
        s = 1 if bit(c, a) == 1 else -1
        e = []
        for d in range(dim):
            if d == a:
                e.append((0, 0))
            else:
                e.append((1, 0) if bit(c, d) == 1 else (0, 1))  # (power of t, power of 1-t)
        return s, e

    def prod_int(terms):
        s = 1
        tot = [[0, 0] for _ in range(dim)]
        for si, e in terms:
            s *= si
            for d in range(dim):
                tot[d][0] += e[d][0]
                tot[d][1] += e[d][1]
        v = Fraction(s)
        for d in range(dim):
            v *= integral_1d(tot[d][0], tot[d][1])
        return v

    pairs = [(i, j) for i in range(dim) for j in range(i)]

    def comp(a, b, i, j):
        return [(Fraction(1, 2), [grad(a, j), grad(b, i)]), (Fraction(-1, 2), [grad(a, i), grad(b, j)])]

    cache = {}

    def M(a, b, c, d):
        key = (a, b, c, d)
        if key not in cache:
            v = Fraction(0)
            for (i, j) in pairs:
                for c1, t1 in comp(a, b, i, j):
                    for c2, t2 in comp(c, d, i, j):
                        v += 2 * c1 * c2 * prod_int(t1 + t2)
            cache[key] = v
        return cache[key]

    kl = [(k, l) for k in range(C) for l in range(k + 1, C)]
    WT = [[M(0, j, k, l) for (k, l) in kl] for j in range(1, C)]
    WA = [[M(0, j, 0, k) for k in range(1, C)] for j in range(1, C)]
    return kl, WT, WA


def basis_integral_float(p, q):
    """_Basis::Integral in double precision, Include/Hat.inl:121-131."""
    def prod(first, last):
        f = 1.0
        for i in range(first, last + 1):
            f *= i
        return f
    if p > q:
        return prod(1, q) / prod(p + 1, p + q + 1)
    return prod(1, p) / prod(q + 1, p + q + 1)


def full_stiffness_float(dim):
    C = 1 << dim

    def bit(c, d):
        return (c >> (dim - 1 - d)) & 1

    def deriv(c, a):
        # (coefficient, p[d], q[d], dd[d]) of d/dt_a of the hat of corner c
        p = [1 if bit(c, d) else 0 for d in range(dim)]
        q = [0 if bit(c, d) else 1 for d in range(dim)]
        dd = [0] * dim
        coef = 1.0
        if p[a]:
            coef *= p[a]; p[a] -= 1; dd[a] = 1
        elif q[a]:
            coef *= -q[a]; q[a] -= 1; dd[a] = 1
        return coef, p, q, dd

    fs = [[[[0.0] * dim for _ in range(dim)] for _ in range(C)] for _ in range(C)]
    for i1 in range(C):
        for i2 in range(C):
            for d1 in range(dim):
                for d2 in range(dim):
                    c1, p1, q1, e1 = deriv(i1, d1)
                    c2, p2, q2, e2 = deriv(i2, d2)
                    I = c1 * c2
                    for d in range(dim):
                        I *= basis_integral_float(p1[d] + p2[d], q1[d] + q2[d]) * 1.0 / 1.0
                    fs[i1][i2][d2][d1] = 0.0 + I
    return fs


def fmt(v):
    return repr(float(v))


def main():
    out = []
    out.append("// Generated by gen_wedge_tables.py -- do not edit.")
    out.append("// Corner-0 wedge element tables at unit resolution; see the generator docstring.")
    out.append("#pragma once")
    for dim in (3, 4):
        kl, WT, WA = tables(dim)
        C = 1 << dim
        P = len(kl)
        out.append(f"namespace wedge{dim} {{")
        out.append(f"constexpr int C = {C};")
        out.append(f"constexpr int P = {P};")
        out.append(f"constexpr int PK[{P}] = {{{', '.join(str(k) for k, _ in kl)}}};")
        out.append(f"constexpr int PL[{P}] = {{{', '.join(str(l) for _, l in kl)}}};")
        out.append(f"constexpr double WT[{C - 1}][{P}] = {{")
        for row in WT:
            out.append("  {" + ", ".join(fmt(v) for v in row) + "},")
        out.append("};")
        out.append(f"constexpr double WA[{C - 1}][{C - 1}] = {{")
        for row in WA:
            out.append("  {" + ", ".join(fmt(v) for v in row) + "},")
        out.append("};")
        fs = full_stiffness_float(dim)
        out.append(f"constexpr double FS[{C}][{C}][{dim}][{dim}] = {{")
        for i1 in range(C):
            out.append("  {" + ", ".join("{" + ", ".join("{" + ", ".join(repr(v) for v in row) + "}" for row in fs[i1][i2]) + "}" for i2 in range(C)) + "},")
        out.append("};")
        out.append(f"}} // namespace wedge{dim}")
    print("\n".join(out))


if __name__ == "__main__":
    main()
