"""ctypes binding of oracle/liboracle.so (TEST INFRASTRUCTURE ONLY).

Mirrors the reference's suqa:: API names on a CPU state; see
oracle/suqa_oracle.h for the reference file:line each call follows.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

U32 = C.c_uint32
PU32 = C.POINTER(C.c_uint32)
PD = C.POINTER(C.c_double)


class _State(C.Structure):
    _fields_ = [("nq", U32), ("N", C.c_uint64), ("data", PD), ("re", PD), ("im", PD),
                ("gc_mask", U32), ("rphase_filled", C.c_int), ("rphase_len", U32),
                ("rphase_m", C.c_double * 65)]


class _Pcg(C.Structure):
    _fields_ = [("state", C.c_uint64), ("inc", C.c_uint64)]


class _Qms(C.Structure):
    _fields_ = [("st", C.POINTER(_State)), ("rng", _Pcg), ("seed", C.c_uint64), ("draws", C.c_uint64),
                ("syst_qbits", U32), ("ene_qbits", U32), ("ene_levels", U32), ("nqubits", U32),
                ("max_reverse_attempts", U32),
                ("t_phase_estimation", C.c_double), ("t_PE_factor", C.c_double), ("t_PE_shift", C.c_double),
                ("n_phase_estimation", C.c_int), ("gCi", U32), ("c_acc", U32),
                ("bm_syst", U32 * 3), ("bm_enes", U32 * 32), ("bm_acc", U32),
                ("W_mask", U32), ("W_mask_Enew", U32),
                ("two_registers", U32), ("bm_enes_old", U32 * 32), ("W_mask_Eold", U32), ("compare_draws", U32),
                ("c_factor", C.c_double),
                ("curr_E_old", U32), ("curr_E_old_d", C.c_double), ("record_reverse", C.c_int),
                ("E_measures", PD), ("X_measures", PD), ("n_meas", C.c_size_t), ("cap_meas", C.c_size_t),
                ("reverse_counters", PU32), ("n_rev", C.c_size_t), ("cap_rev", C.c_size_t)]


class Gate(C.Structure):
    _fields_ = [("kind", C.c_int32), ("target", U32), ("nctrl", U32), ("ctrl", U32 * 3), ("angle", C.c_double)]


class PauliTerm(C.Structure):
    _fields_ = [("w", U32), ("qs", U32 * 4), ("types", U32 * 4), ("coef", C.c_double)]


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def _load():
    if not os.path.exists(LIB_PATH):
        build()
    lib = C.CDLL(LIB_PATH)
    PS = C.POINTER(_State)
    sig = {
        "so_create": (PS, [U32]), "so_destroy": (None, [PS]), "so_init_state": (None, [PS]),
        "so_vnorm": (C.c_double, [PS]), "so_vnormalize": (None, [PS]),
        "so_activate_gc_mask": (None, [PS, PU32, U32]), "so_deactivate_gc_mask": (None, [PS, PU32, U32]),
        "so_apply_x": (None, [PS, U32]), "so_apply_y": (None, [PS, U32]), "so_apply_h": (None, [PS, U32]),
        "so_apply_u1": (None, [PS, U32, U32, C.c_double]),
        "so_apply_cx": (None, [PS, U32, U32, U32]),
        "so_apply_mcx": (None, [PS, PU32, PU32, U32, U32]),
        "so_apply_cu1": (None, [PS, U32, U32, C.c_double, U32]),
        "so_apply_mcu1": (None, [PS, PU32, PU32, U32, U32, C.c_double]),
        "so_apply_swap": (None, [PS, U32, U32]),
        "so_apply_phase_list": (C.c_int, [PS, U32, U32, PD, U32]),
        "so_apply_pauli_TP_rotation": (C.c_int, [PS, PU32, PU32, U32, C.c_double]),
        "so_apply_pauli_general": (C.c_int, [PS, PU32, PU32, U32, C.c_double]),
        "so_apply_qft": (None, [PS, PU32, U32]), "so_apply_qft_inverse": (None, [PS, PU32, U32]),
        "so_prob1": (C.c_double, [PS, U32]), "so_measure_qbit": (U32, [PS, U32, C.c_double]),
        "so_apply_reset": (None, [PS, U32, C.c_double]),
        "so_prob_filter": (C.c_double, [PS, PU32, PU32, U32]),
        "so_pcg32_seed": (None, [C.POINTER(_Pcg), C.c_uint64, C.c_uint64]),
        "so_pcg32_next": (U32, [C.POINTER(_Pcg)]), "so_pcg32_doub": (C.c_double, [C.POINTER(_Pcg)]),
        "so_pcg32_randint": (C.c_int, [C.POINTER(_Pcg), C.c_int, C.c_int]),
        "so_qms_create": (C.POINTER(_Qms), [C.c_double, U32, U32, C.c_uint64, C.c_double, C.c_double, C.c_int, C.c_int]),
        "so_qms_create2": (C.POINTER(_Qms), [C.c_double, U32, U32, C.c_uint64, C.c_double, C.c_double, C.c_int, C.c_int,
                                             C.c_int]),
        "so_qms_apply_W": (None, [C.POINTER(_Qms)]),
        "so_qms_destroy": (None, [C.POINTER(_Qms)]),
        "so_qms_init_state": (None, [C.POINTER(_Qms)]),
        "so_qms_metro_step": (C.c_int, [C.POINTER(_Qms), C.c_int]),
        "so_qms_run": (U32, [C.POINTER(_Qms), U32, U32, C.c_int, C.POINTER(C.c_int)]),
        "so_circuit_generate": (U32, [U32, U32, C.c_uint64, C.POINTER(Gate), PU32, PD]),
        "so_circuit_apply": (None, [PS, C.POINTER(Gate)]),
        "so_trotter_terms_generate": (None, [U32, U32, C.c_uint64, C.POINTER(PauliTerm)]),
        "so_trotter_step": (None, [PS, C.POINTER(PauliTerm), U32, C.c_double]),
        "so_qpe_trotter": (None, [PS, U32, U32, C.POINTER(PauliTerm), U32, C.c_double, U32]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    return lib


_lib = None


def lib():
    global _lib
    if _lib is None:
        _lib = _load()
    return _lib


def _u32(xs):
    a = (U32 * max(1, len(xs)))(*[int(x) for x in xs])
    return a


class Pcg:
    """pcg32 + the reference `pcg` wrapper (Rand.hpp:249-291)."""

    def __init__(self, seed):
        self.s = _Pcg()
        lib().so_pcg32_seed(C.byref(self.s), seed, 1)

    def next(self):
        return lib().so_pcg32_next(C.byref(self.s))

    def doub(self):
        return lib().so_pcg32_doub(C.byref(self.s))

    def randint(self, a, b):
        return lib().so_pcg32_randint(C.byref(self.s), a, b)


class State:
    """One oracle state vector (suqa::state + gc_mask + QFT table)."""

    def __init__(self, nq):
        self.p = lib().so_create(nq)
        if not self.p:
            raise MemoryError("oracle state allocation failed")
        self.nq = nq
        self.N = 1 << nq

    def __del__(self):
        if getattr(self, "p", None):
            lib().so_destroy(self.p)
            self.p = None

    # state access -------------------------------------------------------
    def re(self):
        return np.ctypeslib.as_array(self.p.contents.re, shape=(self.N,))

    def im(self):
        return np.ctypeslib.as_array(self.p.contents.im, shape=(self.N,))

    def get(self):
        return self.re().copy(), self.im().copy()

    def set(self, re, im):
        self.re()[:] = re
        self.im()[:] = im

    @property
    def gc_mask(self):
        return self.p.contents.gc_mask

    # API ----------------------------------------------------------------
    def init_state(self):
        lib().so_init_state(self.p)

    def vnorm(self):
        return lib().so_vnorm(self.p)

    def vnormalize(self):
        lib().so_vnormalize(self.p)

    def activate_gc_mask(self, qs):
        lib().so_activate_gc_mask(self.p, _u32(qs), len(qs))

    def deactivate_gc_mask(self, qs):
        lib().so_deactivate_gc_mask(self.p, _u32(qs), len(qs))

    def x(self, q):
        lib().so_apply_x(self.p, q)

    def y(self, q):
        lib().so_apply_y(self.p, q)

    def h(self, q):
        lib().so_apply_h(self.p, q)

    def u1(self, q, phase, q_mask=1):
        lib().so_apply_u1(self.p, q, q_mask, phase)

    def cx(self, qc, qt, q_mask=1):
        lib().so_apply_cx(self.p, qc, qt, q_mask)

    def mcx(self, ctrls, qt, q_mask=None):
        lib().so_apply_mcx(self.p, _u32(ctrls), None if q_mask is None else _u32(q_mask), len(ctrls), qt)

    def cu1(self, qc, qt, phase, q_mask=1):
        lib().so_apply_cu1(self.p, qc, qt, phase, q_mask)

    def mcu1(self, ctrls, qt, phase, q_mask=None):
        lib().so_apply_mcu1(self.p, _u32(ctrls), None if q_mask is None else _u32(q_mask), len(ctrls), qt, phase)

    def swap(self, a, b):
        lib().so_apply_swap(self.p, a, b)

    def phase_list(self, q0, q_size, phases):
        arr = (C.c_double * len(phases))(*phases)
        return lib().so_apply_phase_list(self.p, q0, q_size, arr, len(phases))

    def pauli(self, qs, types, phase):
        return lib().so_apply_pauli_TP_rotation(self.p, _u32(qs), _u32(types), len(qs), phase)

    def pauli_general(self, qs, types, phase):
        return lib().so_apply_pauli_general(self.p, _u32(qs), _u32(types), len(qs), phase)

    def qft(self, qs):
        lib().so_apply_qft(self.p, _u32(qs), len(qs))

    def qft_inverse(self, qs):
        lib().so_apply_qft_inverse(self.p, _u32(qs), len(qs))

    def prob1(self, q):
        return lib().so_prob1(self.p, q)

    def measure(self, q, rdoub):
        return lib().so_measure_qbit(self.p, q, rdoub)

    def reset(self, q, rdoub):
        lib().so_apply_reset(self.p, q, rdoub)

    def prob_filter(self, qs, q_mask):
        return lib().so_prob_filter(self.p, _u32(qs), _u32(q_mask), len(qs))

    def apply_gate(self, g):
        lib().so_circuit_apply(self.p, C.byref(g))


class Qms:
    """One QMS chain on the oracle (qms.cuh + system.cu + qms.cu loop)."""

    def __init__(self, beta, r, seed, ene_min=0.0, ene_max=1.0, pe_steps=10, max_reverse=100,
                 record_reverse=False, two_registers=False, compare_draws=False):
        """two_registers: E_old in its own register; compare_draws: one register
        with the draws of the reference's "comparison with 2enereg" lines."""
        mode = 1 if two_registers else (2 if compare_draws else 0)
        self.p = lib().so_qms_create2(beta, r, max_reverse, seed, ene_min, ene_max, pe_steps,
                                      1 if record_reverse else 0, mode)
        if not self.p:
            raise ValueError("bad QMS parameters")

    def __del__(self):
        if getattr(self, "p", None):
            lib().so_qms_destroy(self.p)
            self.p = None

    def run(self, steps, reset_each=1, thermalization=100):
        rets = (C.c_int * steps)()
        acc = lib().so_qms_run(self.p, steps, reset_each, thermalization, rets)
        return acc, list(rets)

    @property
    def fields(self):
        return self.p.contents

    def measures(self):
        q = self.p.contents
        n = q.n_meas
        E = [q.E_measures[k] for k in range(n)]
        X = [q.X_measures[k] for k in range(n)]
        return E, X

    def reverse_counters(self):
        q = self.p.contents
        return [q.reverse_counters[k] for k in range(q.n_rev)]

    def rng_state(self):
        return self.p.contents.rng.state

    def set_state(self, re, im):
        st = self.p.contents.st.contents
        np.ctypeslib.as_array(st.re, shape=(st.N,))[:] = re
        np.ctypeslib.as_array(st.im, shape=(st.N,))[:] = im

    def apply_W(self):
        lib().so_qms_apply_W(self.p)

    def state(self):
        st = self.p.contents.st.contents
        N = st.N
        return (np.ctypeslib.as_array(st.re, shape=(N,)).copy(),
                np.ctypeslib.as_array(st.im, shape=(N,)).copy())


def circuit(n, layers, seed):
    """Gate list + measurement plan of the random-circuit workload."""
    gates = (Gate * (n * layers))()
    nm = min(8, n)
    mq = (U32 * nm)()
    mr = (C.c_double * nm)()
    ng = lib().so_circuit_generate(n, layers, seed, gates, mq, mr)
    return gates, ng, list(mq), list(mr)


def trotter_terms(nsys, nterms, seed):
    terms = (PauliTerm * nterms)()
    lib().so_trotter_terms_generate(nsys, nterms, seed, terms)
    return terms


def qpe_trotter(st, nsys, nreg, terms, dt, steps):
    """so_qpe_trotter: H on the register, controlled Trotter powers, inverse QFT."""
    lib().so_qpe_trotter(st.p, nsys, nreg, terms, len(terms), dt, steps)

