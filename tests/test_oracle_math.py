"""Pins the oracle where the reference holds no vector: closed forms built
independently in numpy (Kronecker products of textbook matrices), unitarity
invariants and the DFT."""
import itertools
import math

import numpy as np
import pytest

from oracle import pyoracle as po

I2 = np.eye(2, dtype=complex)
PX = np.array([[0, 1], [1, 0]], dtype=complex)
PY = np.array([[0, -1j], [1j, 0]], dtype=complex)
PZ = np.array([[1, 0], [0, -1]], dtype=complex)
PAULI = {0: I2, 1: PX, 2: PY, 3: PZ}


def op_on(n, qs, mats):
    """Dense operator acting with mats[k] on qubit qs[k] (qubit q = bit q of the index)."""
    full = np.array([[1.0 + 0j]])
    for b in reversed(range(n)):
        m = I2
        for q, mm in zip(qs, mats):
            if q == b:
                m = mm
        full = np.kron(full, m)
    return full


def controlled(n, U, gc_qubits):
    """U restricted to the subspace where all gc_qubits read 1, identity elsewhere."""
    N = 1 << n
    mask = sum(1 << q for q in gc_qubits)
    sel = np.array([(i & mask) == mask for i in range(N)])
    out = np.eye(N, dtype=complex)
    out[np.ix_(sel, sel)] = U[np.ix_(sel, sel)]
    return out


def rand_state(st, seed):
    rng = np.random.default_rng(seed)
    v = rng.normal(size=st.N) + 1j * rng.normal(size=st.N)
    v /= np.linalg.norm(v)
    st.set(v.real, v.imag)
    return v


def get(st):
    re, im = st.get()
    return re + 1j * im


N5 = 5


@pytest.mark.parametrize("seed", [1, 2])
def test_pauli_general_all_strings_vs_dense(seed):
    """exp(i phi P) = cos(phi) I + i sin(phi) P for every string of weight 1..4."""
    rng = np.random.default_rng(seed)
    st = po.State(N5)
    cases = 0
    for w in (1, 2, 3, 4):
        for qs in itertools.permutations(range(N5), w):
            if rng.random() > 0.15:
                continue
            types = [int(t) for t in rng.integers(1, 4, size=w)]
            phi = float(rng.uniform(-3, 3))
            for gc in ([], [q for q in range(N5) if q not in qs][:1]):
                v = rand_state(st, cases)
                P = op_on(N5, qs, [PAULI[t] for t in types])
                U = controlled(N5, math.cos(phi) * np.eye(1 << N5) + 1j * math.sin(phi) * P, gc)
                if gc:
                    st.activate_gc_mask(gc)
                assert st.pauli_general(list(qs), types, phi) == 0
                if gc:
                    st.deactivate_gc_mask(gc)
                np.testing.assert_allclose(get(st), U @ v, atol=1e-13, rtol=0)
                cases += 1
    assert cases > 40


def test_reference_pauli_cases_vs_dense():
    """The reference's own special cases (weights 1, 2, Z(XX|YY|ZZ)) equal the closed form."""
    st = po.State(N5)
    k = 0
    for w in (1, 2, 3):
        for qs in itertools.permutations(range(N5), w):
            for types in itertools.product((1, 2, 3), repeat=w):
                if k % 7:
                    k += 1
                    continue
                k += 1
                v = rand_state(st, k)
                phi = 0.3 + 0.01 * k
                rc = st.pauli(list(qs), list(types), phi)
                supported = w < 3 or (sorted(types)[2] == 3 and sorted(types)[0] == sorted(types)[1])
                assert (rc == 0) == supported
                if not supported:
                    continue
                eff = list(types)
                if w == 2 and sorted(types) == [1, 2]:
                    # suqa_cpu.hpp:431-433 / suqa_kernels.cuh (xy): the {X,Y} pair
                    # rotates about Y on the X-labelled qubit and X on the Y-labelled one.
                    eff = [3 - t for t in types]
                P = op_on(N5, qs, [PAULI[t] for t in eff])
                np.testing.assert_allclose(get(st), (math.cos(phi) * np.eye(32) + 1j * math.sin(phi) * P) @ v,
                                           atol=1e-13, rtol=0)


def test_single_and_controlled_gates_vs_dense():
    st = po.State(N5)
    H = np.array([[1, 1], [1, -1]]) / math.sqrt(2)
    for q in range(N5):
        for gc in ([], [(q + 1) % N5], [(q + 2) % N5, (q + 3) % N5]):
            for name, M, fn in [("x", PX, lambda: st.x(q)), ("y", PY, lambda: st.y(q)),
                                ("h", H, lambda: st.h(q)),
                                ("u1", np.diag([1, np.exp(0.7j)]), lambda: st.u1(q, 0.7)),
                                ("u1m0", np.diag([np.exp(0.7j), 1]), lambda: st.u1(q, 0.7, 0))]:
                v = rand_state(st, q)
                if gc:
                    st.activate_gc_mask(gc)
                fn()
                if gc:
                    st.deactivate_gc_mask(gc)
                U = controlled(N5, op_on(N5, [q], [M]), gc)
                np.testing.assert_allclose(get(st), U @ v, atol=1e-14, rtol=0, err_msg=name)


def test_multi_controlled_vs_dense():
    st = po.State(N5)
    N = 1 << N5
    for (ctrls, qm, t) in [([0], [1], 2), ([4], [0], 1), ([0, 2], [1, 0], 4), ([1, 2, 3], [1, 1, 1], 0)]:
        v = rand_state(st, t)
        st.mcx(ctrls, t, qm)
        U = np.zeros((N, N), dtype=complex)
        for i in range(N):
            ok = all(((i >> c) & 1) == m for c, m in zip(ctrls, qm))
            U[i ^ (1 << t) if ok else i, i] = 1
        np.testing.assert_allclose(get(st), U @ v, atol=0, rtol=0)
        v = rand_state(st, t + 10)
        st.mcu1(ctrls, t, 1.1, qm)
        d = np.array([np.exp(1.1j) if (all(((i >> c) & 1) == m for c, m in zip(ctrls, qm)) and (i >> t) & 1) else 1
                      for i in range(N)])
        np.testing.assert_allclose(get(st), d * v, atol=1e-15, rtol=0)
    v = rand_state(st, 99)
    st.swap(1, 3)
    perm = [i ^ (((i >> 1) & 1 ^ (i >> 3) & 1) * 0b1010) for i in range(N)]
    np.testing.assert_allclose(get(st)[perm], v, atol=0, rtol=0)


def test_qft_is_bit_reversed_dft():
    """apply_qft (no final swaps) maps |x> to sum_y e^{2 pi i x y/M} |rev(y)> / sqrt(M)."""
    st = po.State(N5)
    reg = [0, 2, 3]   # register bit k = qubit reg[k]
    m = len(reg)
    M = 1 << m
    st.qft(list(range(N5)))   # fill the one-shot phase table with the largest size
    for x in range(M):
        st.init_state()
        for k in range(m):
            if (x >> k) & 1:
                st.x(reg[k])
        st.qft(reg)
        amp = get(st)
        for y in range(M):
            ry = sum(((y >> k) & 1) << (m - 1 - k) for k in range(m))
            idx = sum(((ry >> k) & 1) << reg[k] for k in range(m))
            assert abs(amp[idx] - np.exp(2j * math.pi * x * y / M) / math.sqrt(M)) < 1e-14
        st.qft_inverse(reg)
        e = np.zeros(st.N, dtype=complex)
        e[sum(((x >> k) & 1) << reg[k] for k in range(m))] = 1
        np.testing.assert_allclose(get(st), e, atol=1e-14)


def test_measure_born_rule_and_collapse():
    st = po.State(N5)
    for q in range(N5):
        for rd in (0.05, 0.5, 0.95):
            v = rand_state(st, q)
            p1 = float(np.sum(np.abs(v[[i for i in range(32) if (i >> q) & 1]]) ** 2))
            assert abs(st.prob1(q) - p1) < 1e-15
            c = st.measure(q, rd)
            assert c == int(rd < p1)
            keep = np.array([((i >> q) & 1) == c for i in range(32)])
            w = np.where(keep, v, 0)
            w /= np.linalg.norm(w)
            np.testing.assert_allclose(get(st), w, atol=1e-15)
            assert abs(st.vnorm() - 1) < 1e-15


def test_phase_list_intent():
    st = po.State(N5)
    phases = [0.1 * k - 0.3 for k in range(8)]
    for gc in ([], [0], [0, 4]):
        v = rand_state(st, 3)
        if gc:
            st.activate_gc_mask(gc)
        assert st.phase_list(1, 3, phases) == 0
        if gc:
            st.deactivate_gc_mask(gc)
        mask = sum(1 << q for q in gc)
        d = np.array([np.exp(1j * phases[(i >> 1) & 7]) if (i & mask) == mask else 1 for i in range(32)])
        np.testing.assert_allclose(get(st), d * v, atol=1e-15)
    assert st.phase_list(0, 2, [0.0] * 3) != 0


def test_qms_W_closed_form():
    """apply_W rotates (|0>_acc, |1>_acc) by [[fs2, fs1],[fs1, -fs2]] with fs1 = exp(-dE/2) for dE>0."""
    q = po.Qms(1.3, 3, 5, -1.0, 3.0)
    f = q.fields
    st = f.st.contents
    n = f.nqubits
    N = 1 << n
    rng = np.random.default_rng(0)
    v = rng.normal(size=N) + 1j * rng.normal(size=N)
    re = np.ctypeslib.as_array(st.re, shape=(N,))
    im = np.ctypeslib.as_array(st.im, shape=(N,))
    for Eold in (0, 3, 7):
        re[:] = v.real
        im[:] = v.imag
        f.curr_E_old = Eold
        po.lib().so_qms_apply_W.argtypes = [po.C.POINTER(po._Qms)]
        po.lib().so_qms_apply_W(q.p)
        w = v.copy()
        for i in range(N // 2):
            Enew = (i >> 3) & 7
            dE = Enew * f.c_factor - Eold * f.c_factor
            fs1, fs2 = (math.exp(-dE / 2), math.sqrt(1 - math.exp(-dE))) if dE > 0 else (1.0, 0.0)
            a, b = v[i], v[i + N // 2]
            w[i], w[i + N // 2] = fs2 * a + fs1 * b, fs1 * a - fs2 * b
        np.testing.assert_allclose(re + 1j * im, w, atol=1e-14)


def test_qms_first_readout_is_the_prepared_eigenvalue():
    """init_state prepares an E=-1 eigenstate; with E in [-1,3] the QPE reads k=0 with certainty."""
    q = po.Qms(1.0, 4, 42, -1.0, 3.0)
    po.lib().so_qms_init_state(q.p)
    f = q.fields
    st = po.State.__new__(po.State)
    po.lib().so_qms_apply_Phi.argtypes = [po.C.POINTER(po._Qms)]
    po.lib().so_qms_apply_Phi(q.p)
    re, im = q.state()
    p = re ** 2 + im ** 2
    reg = sum(1 << b for b in f.bm_enes[:4])
    p0 = p[[i for i in range(len(p)) if (i & reg) == 0]].sum()
    assert abs(p0 - 1) < 1e-12
    del st


def test_triangle_spectrum():
    """H = X0X1 + X1X2 + X0X2 has eigenvalues {3 (x2), -1 (x6)}: evolution(t) = exp(-iHt)."""
    st = po.State(3)
    H = sum(op_on(3, [a, b], [PX, PX]) for a, b in [(0, 1), (1, 2), (0, 2)])
    w = np.linalg.eigvalsh(H)
    assert np.allclose(sorted(w), [-1] * 6 + [3] * 2)
    t = 0.37
    U = np.zeros((8, 8), dtype=complex)
    for k in range(8):
        st.init_state()
        e = np.zeros(8)
        e[k] = 1
        st.set(e, np.zeros(8))
        for iii in range(3):
            st.pauli([(iii + 1) % 3, (iii + 2) % 3], [1, 1], -t)
        U[:, k] = get(st)
    ev, V = np.linalg.eigh(H)
    np.testing.assert_allclose(U, V @ np.diag(np.exp(-1j * ev * t)) @ V.conj().T, atol=1e-14)
