"""FP64 CPU restatement of the reference's SCF DIIS (numpy).

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of the device DIIS
(rerdmft_amd/csrc/rr_diis.hip); never by the library.

Follows src/Utils/DIIS.cpp:
  commutator_error   e = F (P S) - (S P) F                                   (DIIS.cpp:51-58)
  Diis.extrapolate   newest-first history of at most max_vectors (error, value) pairs; B_ij =
                     Re sum conj(e_i) e_j; for m = total .. 2: scale = max_i B_ii (zero -> return
                     the value), solve [B/scale -1; -1 0][w; lambda] = [0; -1] by Gaussian
                     elimination with partial pivoting, rejecting the system when min|pivot| /
                     max|pivot| <= 1e-14 (then the oldest pair is left out); result sum_i w_i V_i
                     (DIIS.cpp:80-135, solver :18-45).

Pins (tests/test_oracle_diis.py): the commutator vanishes at self-consistency (P from the
generalized eigenvectors of F) and is antisymmetric for symmetric inputs, and a planted P S F in place
of S P F fails the first check; two-vector weights equal the closed form
w_1 = (B_22 - B_12) / (B_11 - 2 B_12 + B_22); a linear residual e_i = V_i - V* is extrapolated to V*
exactly; three identical errors are singular at every m and return the value unchanged.
"""
import numpy as np


def commutator_error(fock, density, overlap, planted=False):
    if planted:  # deliberate mistake for the pinning test: P S F instead of S P F
        return fock @ (density @ overlap) - (density @ overlap) @ fock
    return fock @ (density @ overlap) - (overlap @ density) @ fock


def _solve_dense(m, rhs, tol):
    m = np.array(m, dtype=np.float64)
    rhs = np.array(rhs, dtype=np.float64)
    n = rhs.size
    piv = np.zeros(n)
    for k in range(n):
        best = k + int(np.argmax(np.abs(m[k:, k])))
        if best != k:
            m[[k, best]] = m[[best, k]]
            rhs[[k, best]] = rhs[[best, k]]
        piv[k] = abs(m[k, k])
        if piv[k] == 0.0:
            return None
        for i in range(k + 1, n):
            f = m[i, k] / m[k, k]
            m[i, k:] -= f * m[k, k:]
            rhs[i] -= f * rhs[k]
    if not piv.min() > tol * piv.max():
        return None
    x = np.zeros(n)
    for i in range(n - 1, -1, -1):
        x[i] = (rhs[i] - m[i, i + 1:] @ x[i + 1:]) / m[i, i]
    return x


class Diis:
    def __init__(self, max_vectors):
        self.max_vectors = max_vectors
        self.errors, self.values = [], []
        self.last_weights = []

    def extrapolate(self, error, value):
        self.last_weights = []
        if self.max_vectors < 2:
            return np.array(value)
        self.errors.insert(0, np.array(error))
        self.values.insert(0, np.array(value))
        del self.errors[self.max_vectors:]
        del self.values[self.max_vectors:]
        total = len(self.errors)
        b = np.array([[np.real(np.vdot(ei, ej)) for ej in self.errors] for ei in self.errors])
        for m in range(total, 1, -1):
            scale = max(b[i, i] for i in range(m))
            if not scale > 0.0:
                return np.array(value)
            a = np.zeros((m + 1, m + 1))
            a[:m, :m] = b[:m, :m] / scale
            a[:m, m] = a[m, :m] = -1.0
            rhs = np.zeros(m + 1)
            rhs[m] = -1.0
            x = _solve_dense(a, rhs, 1e-14)
            if x is None:
                continue
            self.last_weights = list(x[:m])
            return sum(x[i] * self.values[i] for i in range(m))
        return np.array(value)
