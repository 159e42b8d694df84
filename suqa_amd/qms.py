"""ctypes binding of include/suqa_qms.h (suqa_amd/lib/libsuqa.so): the QMS
driver loop on the plain FP64 GPU path, and the batched-chain kernel."""
import ctypes as C
import os

import numpy as np

from suqa_amd import capi

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libsuqa.so")
DRIVER = os.path.join(HERE, "bin", "suqa_qms")


class Params(C.Structure):
    _fields_ = [("beta", C.c_double), ("metro_steps", C.c_uint32), ("reset_each", C.c_uint32),
                ("ene_qbits", C.c_uint32), ("max_reverse", C.c_uint32), ("seed", C.c_uint64),
                ("ene_min", C.c_double), ("ene_max", C.c_double), ("pe_steps", C.c_int32),
                ("thermalization", C.c_int32), ("record_reverse", C.c_int32), ("two_registers", C.c_int32)]


class Result(C.Structure):
    _fields_ = [("rets", C.POINTER(C.c_int32)), ("rng_after", C.POINTER(C.c_uint64)),
                ("E", C.POINTER(C.c_double)), ("X", C.POINTER(C.c_double)), ("cap_meas", C.c_uint32),
                ("n_meas", C.c_uint32), ("reverse", C.POINTER(C.c_uint32)), ("cap_reverse", C.c_uint32),
                ("n_reverse", C.c_uint32), ("accepted", C.c_uint32), ("seconds", C.c_double),
                ("final_re", C.POINTER(C.c_double)), ("final_im", C.POINTER(C.c_double))]


class ChainSpec(C.Structure):
    _fields_ = [("beta", C.c_double), ("ene_qbits", C.c_uint32), ("seed", C.c_uint64)]


class BatchParams(C.Structure):
    _fields_ = [("metro_steps", C.c_uint32), ("reset_each", C.c_uint32), ("max_reverse", C.c_uint32),
                ("ene_min", C.c_double), ("ene_max", C.c_double), ("pe_steps", C.c_int32),
                ("thermalization", C.c_int32), ("fp32", C.c_uint32), ("two_registers", C.c_uint32)]


class BatchResult(C.Structure):
    _fields_ = [("rets", C.POINTER(C.c_int32)), ("rng_after", C.POINTER(C.c_uint64)),
                ("E", C.POINTER(C.c_double)), ("X", C.POINTER(C.c_double)), ("cap_meas", C.c_uint32),
                ("n_meas", C.POINTER(C.c_uint32)), ("accepted", C.POINTER(C.c_uint32)), ("seconds", C.c_double),
                ("lds_accesses", C.POINTER(C.c_uint64)), ("final_state", C.POINTER(C.c_double))]


SIGNATURES = {
    "suqa_qms_run": (C.c_int, [C.POINTER(Params), C.POINTER(Result)]),
    "suqa_qms_batch_run": (C.c_int, [C.POINTER(ChainSpec), C.c_uint32, C.POINTER(BatchParams),
                                     C.POINTER(BatchResult)]),
}

_lib = None


def lib():
    global _lib
    if _lib is None:
        capi.lib()   # the engine first (libsuqa.so links it)
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"QMS library not built: {LIB_PATH}")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype, f.argtypes = res, args
        _lib = L
    return _lib


def _ptr(a, ct):
    return a.ctypes.data_as(C.POINTER(ct))


def run(beta, steps, reset_each, r, seed, ene_min=0.0, ene_max=1.0, pe_steps=10, thermalization=100,
        max_reverse=100, record_reverse=False, final_state=False, cap_meas=None, two_registers=False):
    """The qms.cu driver loop on the GPU engine; returns a dict of per-step results.
    two_registers: E_old in its own register (3 + 2r + 1 qubits)."""
    p = Params(beta, steps, reset_each, r, max_reverse, seed, ene_min, ene_max, pe_steps, thermalization,
               1 if record_reverse else 0, 1 if two_registers else 0)
    cap = steps if cap_meas is None else cap_meas
    rets = np.zeros(steps, dtype=np.int32)
    rng = np.zeros(steps, dtype=np.uint64)
    E = np.zeros(max(cap, 1))
    X = np.zeros(max(cap, 1))
    rev = np.zeros(steps, dtype=np.uint32)
    res = Result()
    res.rets, res.rng_after = _ptr(rets, C.c_int32), _ptr(rng, C.c_uint64)
    res.E, res.X, res.cap_meas = _ptr(E, C.c_double), _ptr(X, C.c_double), cap
    res.reverse, res.cap_reverse = _ptr(rev, C.c_uint32), steps
    N = 1 << ((2 * r if two_registers else r) + 4)
    fre = np.zeros(N)
    fim = np.zeros(N)
    if final_state:
        res.final_re, res.final_im = _ptr(fre, C.c_double), _ptr(fim, C.c_double)
    rc = lib().suqa_qms_run(C.byref(p), C.byref(res))
    if rc != 0:
        raise capi.SuqaError(rc, "suqa_qms_run failed (see stderr)")
    nm = min(res.n_meas, cap)
    return dict(rets=rets, rng_after=rng, E=E[:nm], X=X[:nm], n_meas=res.n_meas, accepted=res.accepted,
                reverse=rev[:res.n_reverse], seconds=res.seconds,
                final=(fre, fim) if final_state else None)


def batch_run(chains, steps, reset_each=1, max_reverse=100, ene_min=0.0, ene_max=1.0, pe_steps=10,
              thermalization=100, cap_meas=None, rng_after=False, fp32=False, two_registers=False, lds_bytes=False,
              final_state=False):
    """chains: list of (beta, r, seed).  Returns per-chain arrays.  fp32: FP32
    amplitudes with FP64 probabilities (the precision pattern).  two_registers:
    E_old in its own register (2r + 4 qubits per chain, r in 2..6, stored in
    2^(r+4) amplitudes: E_old is definite outside its own phase estimation).
    lds_bytes: also return each chain's algorithmic LDS bytes.  final_state:
    also return each chain's stored amplitudes after its last step (complex128)."""
    n = len(chains)
    specs = (ChainSpec * n)(*[ChainSpec(b, r, s) for (b, r, s) in chains])
    bp = BatchParams(steps, reset_each, max_reverse, ene_min, ene_max, pe_steps, thermalization, 1 if fp32 else 0,
                     1 if two_registers else 0)
    cap = steps if cap_meas is None else cap_meas
    rets = np.zeros((n, steps), dtype=np.int32)
    rng = np.zeros((n, steps), dtype=np.uint64) if rng_after else None
    E = np.zeros((n, max(cap, 1)))
    X = np.zeros((n, max(cap, 1)))
    nm = np.zeros(n, dtype=np.uint32)
    acc = np.zeros(n, dtype=np.uint32)
    res = BatchResult()
    res.rets = _ptr(rets, C.c_int32)
    res.rng_after = _ptr(rng, C.c_uint64) if rng_after else C.POINTER(C.c_uint64)()
    res.E, res.X, res.cap_meas = _ptr(E, C.c_double), _ptr(X, C.c_double), cap
    res.n_meas, res.accepted = _ptr(nm, C.c_uint32), _ptr(acc, C.c_uint32)
    lds = np.zeros(n, dtype=np.uint64) if lds_bytes else None
    res.lds_accesses = _ptr(lds, C.c_uint64) if lds_bytes else C.POINTER(C.c_uint64)()
    sizes = [1 << (r + 4) for (_, r, _) in chains]
    fin = np.zeros(2 * sum(sizes)) if final_state else None
    res.final_state = _ptr(fin, C.c_double) if final_state else C.POINTER(C.c_double)()
    rc = lib().suqa_qms_batch_run(specs, n, C.byref(bp), C.byref(res))
    if rc != 0:
        raise capi.SuqaError(rc, "suqa_qms_batch_run failed (see stderr)")
    states = None
    if final_state:
        z = fin.view(np.complex128)
        offs = np.concatenate(([0], np.cumsum(sizes)))
        states = [z[offs[c]:offs[c + 1]] for c in range(n)]
    return dict(rets=rets, rng_after=rng, E=E, X=X, n_meas=nm, accepted=acc, seconds=res.seconds, lds_bytes=lds,
                final=states)
