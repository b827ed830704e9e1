"""CPU restatement (test infrastructure, the oracle) of the reference's
Include/MG.h: GSRelaxer (:37-77) and MGSolver (:79-195).

Every sum follows the reference's order: a relaxation sums the column of the
updated index in storage order (:68-72); sparse * dense products gather each
output row in increasing column index (the accumulation order of Eigen's
column-major sparse * dense product); the Galerkin products
(P^T M) P (:96-97) sum every entry's terms in increasing inner index, as
Eigen's conservative sparse * sparse product does.  The coarse solve uses
numpy's dense solver where the reference uses Eigen's SimplicialLDLT, so
results agree with the reference to rounding, not bit for bit.
Pure-Python loops: for the small systems of tests/ only.
"""
import numpy as np
import scipy.sparse as sp


def _csc(A):
    A = sp.csc_matrix(A)
    A.sort_indices()
    return A


def spmv(A, x):
    """y = A x, each row summed in increasing column order (starting from 0)."""
    R = sp.csr_matrix(A)
    R.sort_indices()
    y = np.zeros(R.shape[0])
    for i in range(R.shape[0]):
        s = 0.0
        for k in range(R.indptr[i], R.indptr[i + 1]):
            s += float(R.data[k]) * float(x[R.indices[k]])
        y[i] = s
    return y


def spgemm(A, B):
    """C = A B with every entry's terms in increasing inner index k."""
    A, B = _csc(A), _csc(B)
    cols = []
    for j in range(B.shape[1]):
        acc = {}
        order = []
        for q in range(B.indptr[j], B.indptr[j + 1]):
            k, y = B.indices[q], float(B.data[q])
            for p in range(A.indptr[k], A.indptr[k + 1]):
                i, x = A.indices[p], float(A.data[p])
                if i in acc:
                    acc[i] += x * y
                else:
                    acc[i] = x * y
                    order.append(i)
        cols.append(acc)
    rows, cidx, vals = [], [], []
    for j, acc in enumerate(cols):
        for i in sorted(acc):
            rows.append(i), cidx.append(j), vals.append(acc[i])
    return sp.csc_matrix((vals, (rows, cidx)), shape=(A.shape[0], B.shape[1]))


class GSRelaxer:
    def __init__(self, M):
        self.M = _csc(M)
        self.D = self.M.diagonal().copy()

    def __call__(self, b, x, mc):
        M = self.M
        for lst in mc:
            for i in lst:
                value = 0.0
                for k in range(M.indptr[i], M.indptr[i + 1]):
                    value += float(M.data[k]) * float(x[M.indices[k]])
                x[i] += (b[i] - value) / self.D[i]
        return x


class MGSolver:
    def __init__(self, M, Ps):
        self.M = _csc(M)
        self.Ps = [_csc(P) for P in Ps]
        self.Rs = [_csc(P.T) for P in self.Ps]
        p = len(self.Ps)
        self.Ms = [None] * p
        if p:
            self.Ms[-1] = spgemm(spgemm(self.Rs[-1], self.M), self.Ps[-1])
            for i in range(p - 2, -1, -1):
                self.Ms[i] = spgemm(spgemm(self.Rs[i], self.Ms[i + 1]), self.Ps[i])
        self.relax = [GSRelaxer(self.Ms[i]) for i in range(1, p)] + [GSRelaxer(self.M)]
        self.coarse = (self.Ms[0] if p else self.M).toarray()

    def vcycle(self, b, x, iters, mci):
        p = len(self.Ps)
        X = [None] * p
        B = [None] * p
        for l in range(p, 0, -1):
            if l == p:
                for _ in range(iters):
                    self.relax[l - 1](b, x, mci(l))
                B[l - 1] = spmv(self.Rs[l - 1], b - spmv(self.M, x))
            else:
                for _ in range(iters):
                    self.relax[l - 1](B[l], X[l], mci(l))
                B[l - 1] = spmv(self.Rs[l - 1], B[l] - spmv(self.Ms[l], X[l]))
            X[l - 1] = np.zeros(len(B[l - 1]))
        if p:
            X[0] = np.linalg.solve(self.coarse, B[0])
        else:
            x[:] = np.linalg.solve(self.coarse, b)
        for l in range(1, p + 1):
            if l == p:
                x += spmv(self.Ps[l - 1], X[l - 1])
                for _ in range(iters):
                    self.relax[l - 1](b, x, mci(l))
            else:
                X[l] += spmv(self.Ps[l - 1], X[l - 1])
                for _ in range(iters):
                    self.relax[l - 1](B[l], X[l], mci(l))
        return x
