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
        out.append(f"}} // namespace wedge{dim}")
    print("\n".join(out))


if __name__ == "__main__":
    main()
