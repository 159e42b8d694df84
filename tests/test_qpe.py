"""The Trotter + phase-estimation workload on the oracle (no GPU): the QPE
composition (register powers, the no-swap inverse QFT, the LSB-first bin
order) against textbook phase estimation, and the Trotter product against
dense linear algebra (Pauli strings as Kronecker products of the 2x2 Pauli
matrices, eigen-decomposition of the Trotter step)."""
import numpy as np
import pytest

from oracle import pyoracle as po
from suqa_amd import workloads as wl

PAULI = {1: np.array([[0, 1], [1, 0]], complex), 2: np.array([[0, -1j], [1j, 0]]),
         3: np.array([[1, 0], [0, -1]], complex)}


def textbook_qpe(theta, m):
    """P(y) of m-bit phase estimation of the eigenphase e^{i theta}."""
    M = 1 << m
    x = np.arange(M)
    return np.array([abs(np.exp(1j * x * (theta - 2 * np.pi * y / M)).sum() / M) ** 2 for y in range(M)])


def terms_from(spec):
    out = []
    for qs, types, coef in spec:
        t = po.PauliTerm()
        t.w = len(qs)
        for k, (q, ty) in enumerate(zip(qs, types)):
            t.qs[k], t.types[k] = q, ty
        t.coef = coef
        out.append(t)
    return (po.PauliTerm * len(out))(*out)


def oracle_hist(nsys, nreg, terms, dt, steps):
    st = po.State(nsys + nreg)
    po.qpe_trotter(st, nsys, nreg, terms, dt, steps)
    reg = list(range(nsys, nsys + nreg))
    return np.array([st.prob_filter(reg, [(k >> j) & 1 for j in range(nreg)]) for k in range(1 << nreg)])


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_qpe_of_an_eigenstate_is_textbook(seed):
    """Z strings on |0..0>: an eigenstate of every term, eigenphase
    theta = -dt * steps * sum(coef); bin k holds the textbook P(k)."""
    rng = np.random.default_rng(seed)
    nsys, nreg, dt, steps = 5, 4, 0.37, 3
    spec = []
    for _ in range(7):
        w = int(rng.integers(2, 5))
        qs = [int(q) for q in rng.choice(nsys, w, replace=False)]
        spec.append((qs, [3] * w, float(rng.uniform(-1, 1))))
    terms = terms_from(spec)
    theta = -dt * steps * sum(c for _, _, c in spec)
    np.testing.assert_allclose(oracle_hist(nsys, nreg, terms, dt, steps), textbook_qpe(theta, nreg),
                               atol=1e-12, rtol=0)


def dense_pauli(n, qs, types):
    ops = [np.eye(2, dtype=complex)] * n
    for q, t in zip(qs, types):
        ops[q] = PAULI[t]
    m = np.array([[1.0 + 0j]])
    for q in range(n - 1, -1, -1):    # kron(A, B): A on the higher qubit
        m = np.kron(m, ops[q])
    return m


def test_qpe_of_random_terms_vs_eigendecomposition():
    nsys, nreg, dt, steps = 4, 3, 0.05, 4
    terms = po.trotter_terms(nsys, 12, 11)
    U1 = np.eye(1 << nsys, dtype=complex)
    for t in terms:
        P = dense_pauli(nsys, list(t.qs[:t.w]), list(t.types[:t.w]))
        phi = -t.coef * dt
        U1 = (np.cos(phi) * np.eye(1 << nsys) + 1j * np.sin(phi) * P) @ U1
    U = np.linalg.matrix_power(U1, steps)
    ev, V = np.linalg.eig(U)
    psi0 = np.zeros(1 << nsys, complex)
    psi0[0] = 1
    c = np.linalg.solve(V, psi0)
    # U is unitary with (generically) distinct eigenvalues, so V is unitary up to
    # column scaling: weights |c_j|^2 |v_j|^2
    w = np.abs(c) ** 2 * np.linalg.norm(V, axis=0) ** 2
    assert abs(w.sum() - 1) < 1e-10
    want = sum(wj * textbook_qpe(np.angle(e), nreg) for wj, e in zip(w, ev))
    np.testing.assert_allclose(oracle_hist(nsys, nreg, terms, dt, steps), want, atol=1e-10, rtol=0)


def test_trotter_step_ops_match_oracle_terms():
    terms = wl.trotter_terms(24, 96, 11)
    ot = po.trotter_terms(24, 96, 11)
    ops = wl.trotter_step_ops(terms, 0.05, 1 << 27)
    for (qs, types, phase, gc), t in zip(ops, ot):
        assert qs == list(t.qs[:t.w]) and types == list(t.types[:t.w])
        assert phase == -t.coef * 0.05 and gc == 1 << 27
