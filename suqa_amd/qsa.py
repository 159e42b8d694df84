"""ctypes binding of include/suqa_qsa.h (suqa_amd/lib/libsuqa.so): the QSA
driver loop of the reference (qsa.cu) on the plain FP64 GPU path."""
import ctypes as C
import os

import numpy as np

from suqa_amd import capi, qms

HERE = os.path.dirname(os.path.abspath(__file__))
DRIVER = os.path.join(HERE, "bin", "suqa_qsa")


class Params(C.Structure):
    _fields_ = [("beta", C.c_double), ("sampling", C.c_uint32), ("ene_qbits", C.c_uint32),
                ("szegedy_qbits", C.c_uint32), ("seed", C.c_uint64), ("ene_min", C.c_double),
                ("ene_max", C.c_double), ("ene_min_szegedy", C.c_double), ("ene_max_szegedy", C.c_double),
                ("pe_steps", C.c_int32), ("annealing_sequences", C.c_uint32), ("max_annealing_steps", C.c_uint32)]


class Result(C.Structure):
    _fields_ = [("steps", C.POINTER(C.c_uint32)), ("rng_after", C.POINTER(C.c_uint64)), ("cap_steps", C.c_uint32),
                ("n_steps", C.c_uint32), ("E", C.POINTER(C.c_double)), ("X", C.POINTER(C.c_double)),
                ("rejections", C.POINTER(C.c_int32)), ("sample_rng", C.POINTER(C.c_uint64)),
                ("cap_samples", C.c_uint32), ("n_samples", C.c_uint32), ("final_rng", C.c_uint64),
                ("final_re", C.POINTER(C.c_double)), ("final_im", C.POINTER(C.c_double)), ("seconds", C.c_double)]


SIGNATURES = {"suqa_qsa_run": (C.c_int, [C.POINTER(Params), C.POINTER(Result)])}

_lib = None


def lib():
    global _lib
    if _lib is None:
        L = qms.lib()   # libsuqa.so (and the engine it links)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype, f.argtypes = res, args
        _lib = L
    return _lib


def _ptr(a, ct):
    return a.ctypes.data_as(C.POINTER(ct))


def run(beta, sampling, ene_qbits, szegedy_qbits, seed, ene_min=-4.0, ene_max=8.0, ene_min_szegedy=-1.0,
        ene_max_szegedy=3.0, pe_steps=1, annealing_sequences=100, max_annealing_steps=0, cap_steps=4096,
        final_state=False):
    """The qsa.cu driver loop on the GPU engine; returns per-step and per-sample results."""
    p = Params(beta, sampling, ene_qbits, szegedy_qbits, seed, ene_min, ene_max, ene_min_szegedy, ene_max_szegedy,
               pe_steps, annealing_sequences, max_annealing_steps)
    steps = np.zeros((cap_steps, 5), dtype=np.uint32)
    rng = np.zeros(cap_steps, dtype=np.uint64)
    E = np.zeros(max(sampling, 1))
    X = np.zeros(max(sampling, 1))
    rej = np.zeros(max(sampling, 1), dtype=np.int32)
    srng = np.zeros(max(sampling, 1), dtype=np.uint64)
    N = 1 << (3 + ene_qbits + szegedy_qbits + 1)
    fre, fim = np.zeros(N), np.zeros(N)
    r = Result()
    r.steps, r.rng_after, r.cap_steps = _ptr(steps, C.c_uint32), _ptr(rng, C.c_uint64), cap_steps
    r.E, r.X, r.rejections, r.sample_rng = _ptr(E, C.c_double), _ptr(X, C.c_double), _ptr(rej, C.c_int32), \
        _ptr(srng, C.c_uint64)
    r.cap_samples = sampling
    if final_state:
        r.final_re, r.final_im = _ptr(fre, C.c_double), _ptr(fim, C.c_double)
    rc = lib().suqa_qsa_run(C.byref(p), C.byref(r))
    if rc != 0:
        raise capi.SuqaError(rc, "suqa_qsa_run failed (see stderr)")
    ns, nm = min(r.n_steps, cap_steps), min(r.n_samples, sampling)
    return dict(steps=steps[:ns], rng_after=rng[:ns], n_steps=r.n_steps, E=E[:nm], X=X[:nm], rejections=rej[:nm],
                sample_rng=srng[:nm], final_rng=r.final_rng, seconds=r.seconds,
                final=(fre, fim) if final_state else None)
