"""ctypes binding of the C-ABI in include/suqa_hip.h (suqa_amd/lib/libsuqa_hip.so).

`Engine` exposes the reference's suqa:: API names (suqa.cuh:94-176) over one
GPU state; gc_mask is held on the Python side and passed to every gate, as
the reference's global suqa::gc_mask is (suqa.cu:36, 45-61).  There is no
CPU fallback: if the HIP library is missing or fails to load, importing
this module raises.
"""
import ctypes as C
import math
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libsuqa_hip.so")

U32 = C.c_uint32
PU32 = C.POINTER(C.c_uint32)
PD = C.POINTER(C.c_double)
CTX = C.c_void_p

SUQA_OK = 0
SUQA_ERR_ARG = -1
SUQA_ERR_UNSUPPORTED = -2
OUTCOME_SLOTS = 1024   # SUQA_HIP_OUTCOME_SLOTS: the device ring of deferred outcomes

GATE_H, GATE_X, GATE_RX, GATE_RZ, GATE_U1, GATE_CX, GATE_MCX = range(7)


class Gate(C.Structure):
    _fields_ = [("kind", C.c_int32), ("target", U32), ("nctrl", U32), ("ctrl", U32 * 3), ("angle", C.c_double)]


class PauliOp(C.Structure):
    _fields_ = [("w", U32), ("qs", U32 * 8), ("types", U32 * 8), ("gc_mask", U32), ("phase", C.c_double)]


class KStat(C.Structure):
    _fields_ = [("name", C.c_char * 32), ("launches", C.c_uint64), ("total_ms", C.c_double),
                ("alg_bytes", C.c_double), ("flops", C.c_double)]


# name -> (restype, argtypes); every entry point of include/suqa_hip.h
SIGNATURES = {
    "suqa_hip_last_error": (C.c_char_p, []),
    "suqa_hip_setup": (C.c_int, [U32, C.POINTER(CTX)]),
    "suqa_hip_clear": (C.c_int, [CTX]),
    "suqa_hip_allocate_state": (C.c_int, [CTX, U32]),
    "suqa_hip_init_state": (C.c_int, [CTX]),
    "suqa_hip_vnorm": (C.c_int, [CTX, PD]),
    "suqa_hip_vnormalize": (C.c_int, [CTX]),
    "suqa_hip_apply_x": (C.c_int, [CTX, U32, U32]),
    "suqa_hip_apply_y": (C.c_int, [CTX, U32, U32]),
    "suqa_hip_apply_h": (C.c_int, [CTX, U32, U32]),
    "suqa_hip_apply_u1": (C.c_int, [CTX, U32, U32, C.c_double, U32]),
    "suqa_hip_apply_mcx": (C.c_int, [CTX, PU32, PU32, U32, U32, U32]),
    "suqa_hip_apply_mcu1": (C.c_int, [CTX, PU32, PU32, U32, U32, C.c_double, U32]),
    "suqa_hip_apply_swap": (C.c_int, [CTX, U32, U32, U32]),
    "suqa_hip_apply_phase_list": (C.c_int, [CTX, U32, U32, PD, U32, U32]),
    "suqa_hip_apply_pauli_TP_rotation": (C.c_int, [CTX, PU32, PU32, U32, C.c_double, U32]),
    "suqa_hip_apply_pauli_rotation_general": (C.c_int, [CTX, PU32, PU32, U32, C.c_double, U32]),
    "suqa_hip_apply_qft": (C.c_int, [CTX, PU32, U32, U32]),
    "suqa_hip_apply_qft_inverse": (C.c_int, [CTX, PU32, U32, U32]),
    "suqa_hip_measure_qbit": (C.c_int, [CTX, U32, C.c_double, PU32]),
    "suqa_hip_apply_reset": (C.c_int, [CTX, U32, C.c_double, U32]),
    "suqa_hip_measure_qbits": (C.c_int, [CTX, PU32, C.POINTER(C.c_double), U32, PU32]),
    "suqa_hip_apply_resets": (C.c_int, [CTX, PU32, C.POINTER(C.c_double), U32, U32]),
    "suqa_hip_measure_qbits_deferred": (C.c_int, [CTX, PU32, C.POINTER(C.c_double), U32, PU32]),
    "suqa_hip_read_outcomes": (C.c_int, [CTX, U32, U32, PU32]),
    "suqa_hip_apply_x_if": (C.c_int, [CTX, PU32, U32, U32, U32]),
    "suqa_hip_resets_then_x_if": (C.c_int, [CTX, PU32, PD, PU32, U32, U32, U32]),
    "suqa_hip_qms_apply_W_deferred": (C.c_int, [CTX, U32, U32, U32, C.c_double, U32]),
    "suqa_hip_prob_filter": (C.c_int, [CTX, PU32, PU32, U32, PD]),
    "suqa_hip_qms_apply_W": (C.c_int, [CTX, U32, U32, U32, C.c_double, U32]),
    "suqa_hip_qms_apply_W2": (C.c_int, [CTX, U32, U32, U32, U32, U32, C.c_double]),
    "suqa_hip_state_info": (C.c_int, [CTX, PU32, C.POINTER(C.c_uint64), C.POINTER(PD), C.POINTER(PD)]),
    "suqa_hip_download": (C.c_int, [CTX, PD, PD]),
    "suqa_hip_upload": (C.c_int, [CTX, PD, PD]),
    "suqa_hip_synchronize": (C.c_int, [CTX]),
    "suqa_hip_checksum": (C.c_int, [CTX, PD]),
    "suqa_hip_graph_begin": (C.c_int, [CTX]),
    "suqa_hip_graph_end": (C.c_int, [CTX, PU32]),
    "suqa_hip_graph_launch": (C.c_int, [CTX, U32]),
    "suqa_hip_run_gates": (C.c_int, [CTX, C.POINTER(Gate), U32]),
    "suqa_hip_timing_enable": (C.c_int, [CTX, C.c_int]),
    "suqa_hip_timing_reset": (C.c_int, [CTX]),
    "suqa_hip_timing_report": (C.c_int, [CTX, C.POINTER(KStat), U32, PU32]),
    "suqa_hip_copy_kernel": (C.c_int, [CTX, C.c_void_p, C.c_void_p, C.c_uint64]),
    "suqa_hip_device_alloc": (C.c_int, [C.c_uint64, C.POINTER(C.c_void_p)]),
    "suqa_hip_device_free": (C.c_int, [C.c_void_p]),
    "suqa_hip_prob_filter_noalloc": (C.c_int, [PD, PD, U32, PU32, PU32, U32, PD, C.c_void_p]),
    "suqa_hip_run_paulis": (C.c_int, [CTX, C.POINTER(PauliOp), U32]),
    "suqa_hip_set_tile_bits": (C.c_int, [CTX, U32]),
    "suqa_hip_get_tile_bits": (C.c_int, [CTX, PU32]),
    "suqa_hip_set_gate_order": (C.c_int, [CTX, C.c_int]),
    "suqa_hip_autotune_tile_bits": (C.c_int, [CTX, C.POINTER(Gate), U32, PU32, U32, PU32, PD]),
    "suqa_hip_pass_stats": (C.c_int, [CTX, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]),
    # FP32 fused-block engine (extension)
    "suqa_hip_f32_last_error": (C.c_char_p, []),
    "suqa_hip_f32_setup": (C.c_int, [U32, C.POINTER(CTX)]),
    "suqa_hip_f32_clear": (C.c_int, [CTX]),
    "suqa_hip_f32_init_state": (C.c_int, [CTX]),
    "suqa_hip_f32_run_gates": (C.c_int, [CTX, C.POINTER(Gate), U32, U32, PU32]),
    "suqa_hip_f32_run_paulis": (C.c_int, [CTX, C.POINTER(PauliOp), U32, U32, PU32]),
    "suqa_hip_f32_qft_inverse": (C.c_int, [CTX, PU32, U32, U32]),
    "suqa_hip_f32_autotune": (C.c_int, [CTX, C.POINTER(Gate), U32, PU32, PD]),
    "suqa_hip_f32_set_passes": (C.c_int, [CTX, U32]),
    "suqa_hip_f32_get_passes": (C.c_int, [CTX, PU32]),
    "suqa_hip_f32_set_tile_bits": (C.c_int, [CTX, U32]),
    "suqa_hip_f32_get_tile_bits": (C.c_int, [CTX, PU32]),
    "suqa_hip_f32_vnorm": (C.c_int, [CTX, PD]),
    "suqa_hip_f32_prob_filter": (C.c_int, [CTX, PU32, PU32, U32, PD]),
    "suqa_hip_f32_measure_qbit": (C.c_int, [CTX, U32, C.c_double, PU32]),
    "suqa_hip_f32_fidelity": (C.c_int, [CTX, CTX, PD]),
    "suqa_hip_f32_download": (C.c_int, [CTX, C.POINTER(C.c_float), C.POINTER(C.c_float)]),
    "suqa_hip_f32_synchronize": (C.c_int, [CTX]),
    "suqa_hip_f32_timing_enable": (C.c_int, [CTX, C.c_int]),
    "suqa_hip_f32_timing_reset": (C.c_int, [CTX]),
    "suqa_hip_f32_timing_report": (C.c_int, [CTX, C.POINTER(KStat), U32, PU32]),
    "suqa_hip_fuse_plan": (C.c_int, [C.POINTER(Gate), U32, U32, U32, U32, PU32, PU32, PU32, PD]),
}


class SuqaError(RuntimeError):
    """A non-zero status from the C-ABI (the reference raises std::runtime_error)."""

    def __init__(self, code, msg):
        super().__init__(f"suqa_hip error {code}: {msg}")
        self.code = code


def load_library(path=LIB_PATH):
    if not os.path.exists(path):
        raise ImportError(f"HIP engine library not built: {path} (run __graft_entry__.build())")
    lib = C.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    return lib


_lib = None


def lib():
    global _lib
    if _lib is None:
        _lib = load_library()
    return _lib


def _u32(xs):
    return (U32 * max(1, len(xs)))(*[int(x) for x in xs])


def check(rc):
    if rc != SUQA_OK:
        raise SuqaError(rc, lib().suqa_hip_last_error().decode())
    return rc


class Engine:
    """One GPU state vector behind the reference's API names."""

    def __init__(self, nq, tile_bits=None):
        """tile_bits: None keeps the library default (SUQA_HIP_TILE_BITS or 11);
        0 applies one sweep per gate."""
        self.ctx = CTX()
        check(lib().suqa_hip_setup(nq, C.byref(self.ctx)))
        self.nq = nq
        self.N = 1 << nq
        self.gc_mask = 0
        if tile_bits is not None:
            self.set_tile_bits(tile_bits)

    def close(self):
        if self.ctx:
            lib().suqa_hip_clear(self.ctx)
            self.ctx = CTX()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # suqa.cu:45-61
    def activate_gc_mask(self, qs):
        for q in qs:
            self.gc_mask |= 1 << q

    def deactivate_gc_mask(self, qs):
        for q in qs:
            self.gc_mask &= ~(1 << q)

    def allocate_state(self, nq):
        check(lib().suqa_hip_allocate_state(self.ctx, nq))
        self.nq, self.N = nq, 1 << nq

    def init_state(self):
        check(lib().suqa_hip_init_state(self.ctx))

    def vnorm(self):
        v = C.c_double()
        check(lib().suqa_hip_vnorm(self.ctx, C.byref(v)))
        return v.value

    def vnormalize(self):
        check(lib().suqa_hip_vnormalize(self.ctx))

    def x(self, q):
        check(lib().suqa_hip_apply_x(self.ctx, q, self.gc_mask))

    def y(self, q):
        check(lib().suqa_hip_apply_y(self.ctx, q, self.gc_mask))

    def z(self, q):
        self.u1(q, math.pi)

    def s(self, q):
        self.u1(q, math.pi / 2.0)

    def t(self, q):
        self.u1(q, math.pi / 4.0)

    def tdg(self, q):
        self.u1(q, -math.pi / 4.0)

    def h(self, q):
        check(lib().suqa_hip_apply_h(self.ctx, q, self.gc_mask))

    def u1(self, q, phase, q_mask=1):
        check(lib().suqa_hip_apply_u1(self.ctx, q, q_mask, phase, self.gc_mask))

    def cx(self, qc, qt, q_mask=1):
        check(lib().suqa_hip_apply_mcx(self.ctx, _u32([qc]), _u32([q_mask]), 1, qt, self.gc_mask))

    def mcx(self, ctrls, qt, q_mask=None):
        vals = None if q_mask is None else _u32(q_mask)
        check(lib().suqa_hip_apply_mcx(self.ctx, _u32(ctrls), vals, len(ctrls), qt, self.gc_mask))

    def cu1(self, qc, qt, phase, q_mask=1):
        check(lib().suqa_hip_apply_mcu1(self.ctx, _u32([qc]), _u32([q_mask]), 1, qt, phase, self.gc_mask))

    def mcu1(self, ctrls, qt, phase, q_mask=None):
        vals = None if q_mask is None else _u32(q_mask)
        check(lib().suqa_hip_apply_mcu1(self.ctx, _u32(ctrls), vals, len(ctrls), qt, phase, self.gc_mask))

    def swap(self, a, b):
        check(lib().suqa_hip_apply_swap(self.ctx, a, b, self.gc_mask))

    def phase_list(self, q0, q_size, phases):
        arr = (C.c_double * max(1, len(phases)))(*phases)
        return lib().suqa_hip_apply_phase_list(self.ctx, q0, q_size, arr, len(phases), self.gc_mask)

    def pauli(self, qs, types, phase):
        """Returns 0, or SUQA_ERR_UNSUPPORTED where the reference throws."""
        rc = lib().suqa_hip_apply_pauli_TP_rotation(self.ctx, _u32(qs), _u32(types), len(qs), phase,
                                                    self.gc_mask)
        if rc not in (SUQA_OK, SUQA_ERR_UNSUPPORTED):
            check(rc)
        return rc

    def pauli_general(self, qs, types, phase):
        return check(lib().suqa_hip_apply_pauli_rotation_general(self.ctx, _u32(qs), _u32(types), len(qs),
                                                                 phase, self.gc_mask))

    def qft(self, qs):
        check(lib().suqa_hip_apply_qft(self.ctx, _u32(qs), len(qs), self.gc_mask))

    def qft_inverse(self, qs):
        check(lib().suqa_hip_apply_qft_inverse(self.ctx, _u32(qs), len(qs), self.gc_mask))

    def measure(self, q, rdoub):
        c = U32()
        check(lib().suqa_hip_measure_qbit(self.ctx, q, rdoub, C.byref(c)))
        return c.value

    def reset(self, q, rdoub):
        check(lib().suqa_hip_apply_reset(self.ctx, q, rdoub, self.gc_mask))

    def measure_qbits(self, qs, rdoubs):
        cs = (U32 * max(1, len(qs)))()
        rd = (C.c_double * max(1, len(rdoubs)))(*rdoubs)
        check(lib().suqa_hip_measure_qbits(self.ctx, _u32(qs), rd, len(qs), cs))
        return [cs[k] for k in range(len(qs))]

    def resets(self, qs, rdoubs):
        rd = (C.c_double * max(1, len(rdoubs)))(*rdoubs)
        check(lib().suqa_hip_apply_resets(self.ctx, _u32(qs), rd, len(qs), self.gc_mask))

    def measure_qbits_deferred(self, qs, rdoubs):
        """Queued measurements; returns the first outcome slot (outcomes: read_outcomes)."""
        rd = (C.c_double * max(1, len(rdoubs)))(*rdoubs)
        slot = U32()
        check(lib().suqa_hip_measure_qbits_deferred(self.ctx, _u32(qs), rd, len(qs), C.byref(slot)))
        return slot.value

    def read_outcomes(self, slot, n):
        out = (U32 * max(1, n))()
        check(lib().suqa_hip_read_outcomes(self.ctx, slot, n, out))
        return [out[k] for k in range(n)]

    def x_if(self, qs, slot):
        check(lib().suqa_hip_apply_x_if(self.ctx, _u32(qs), len(qs), slot, self.gc_mask))

    def resets_then_x_if(self, qs, rdoubs, xqs, slot):
        """Each reset of qs[k], then X on xqs[k] where outcome slot + k reads 1."""
        assert len(qs) == len(rdoubs) == len(xqs)
        rd = (C.c_double * max(1, len(rdoubs)))(*rdoubs)
        check(lib().suqa_hip_resets_then_x_if(self.ctx, _u32(qs), rd, _u32(xqs), len(qs), slot, self.gc_mask))

    def qms_apply_W_deferred(self, q_acc, w_mask_enew, bm_enes0, c_factor, eold_slot):
        check(lib().suqa_hip_qms_apply_W_deferred(self.ctx, q_acc, w_mask_enew, bm_enes0, c_factor, eold_slot))

    def prob_filter(self, qs, q_mask):
        p = C.c_double()
        check(lib().suqa_hip_prob_filter(self.ctx, _u32(qs), _u32(q_mask), len(qs), C.byref(p)))
        return p.value

    def qms_apply_W(self, q_acc, w_mask_enew, bm_enes0, c_factor, curr_E_old):
        check(lib().suqa_hip_qms_apply_W(self.ctx, q_acc, w_mask_enew, bm_enes0, c_factor, curr_E_old))

    # state access -----------------------------------------------------------
    def get(self):
        re = np.empty(self.N, dtype=np.float64)
        im = np.empty(self.N, dtype=np.float64)
        check(lib().suqa_hip_download(self.ctx, re.ctypes.data_as(PD), im.ctypes.data_as(PD)))
        return re, im

    def set(self, re, im):
        re = np.ascontiguousarray(re, dtype=np.float64)
        im = np.ascontiguousarray(im, dtype=np.float64)
        check(lib().suqa_hip_upload(self.ctx, re.ctypes.data_as(PD), im.ctypes.data_as(PD)))

    def re(self):
        return self.get()[0]

    def im(self):
        return self.get()[1]

    def synchronize(self):
        check(lib().suqa_hip_synchronize(self.ctx))

    def graph_begin(self):
        """Start capturing the context stream (gate calls are recorded, not run)."""
        check(lib().suqa_hip_graph_begin(self.ctx))

    def graph_end(self):
        g = U32()
        check(lib().suqa_hip_graph_end(self.ctx, C.byref(g)))
        return g.value

    def graph_launch(self, g):
        check(lib().suqa_hip_graph_launch(self.ctx, g))

    def checksum(self):
        """(sum re, sum im, sum_i i |psi_i|^2 / N), FP64 tree sums on the device."""
        out = (C.c_double * 3)()
        check(lib().suqa_hip_checksum(self.ctx, out))
        return tuple(out)

    def run_gates(self, gates, n=None, start=0):
        """Applies gates[start:start+n] (a Gate array) asynchronously."""
        n = len(gates) - start if n is None else n
        ptr = C.cast(C.addressof(gates) + start * C.sizeof(Gate), C.POINTER(Gate))
        check(lib().suqa_hip_run_gates(self.ctx, ptr, n))

    def run_paulis(self, ops):
        check(lib().suqa_hip_run_paulis(self.ctx, ops, len(ops)))

    def set_tile_bits(self, bits):
        check(lib().suqa_hip_set_tile_bits(self.ctx, bits))

    def set_gate_order(self, packed):
        """0: call order (bit-identical to one sweep per gate); 1: commuting gates packed."""
        check(lib().suqa_hip_set_gate_order(self.ctx, 1 if packed else 0))

    def tile_bits(self):
        v = U32()
        check(lib().suqa_hip_get_tile_bits(self.ctx, C.byref(v)))
        return v.value

    def autotune_tile_bits(self, gates, candidates=(8, 9, 10, 11, 12)):
        """Times the gate sample at each tile size (changing the state), keeps the fastest."""
        best = U32()
        t = (C.c_double * len(candidates))()
        check(lib().suqa_hip_autotune_tile_bits(self.ctx, gates, len(gates), _u32(candidates), len(candidates),
                                                C.byref(best), t))
        return best.value, {b: t[k] for k, b in enumerate(candidates)}

    def pass_stats(self):
        a, b, c = C.c_uint64(), C.c_uint64(), C.c_uint64()
        check(lib().suqa_hip_pass_stats(self.ctx, C.byref(a), C.byref(b), C.byref(c)))
        return {"passes": a.value, "pass_gates": b.value, "single_sweeps": c.value}

    def apply_gate(self, g):
        self.run_gates((Gate * 1)(g), 1)

    def timing(self, on):
        check(lib().suqa_hip_timing_enable(self.ctx, 1 if on else 0))

    def timing_reset(self):
        check(lib().suqa_hip_timing_reset(self.ctx))

    def timing_report(self):
        arr = (KStat * 64)()
        n = U32()
        check(lib().suqa_hip_timing_report(self.ctx, arr, 64, C.byref(n)))
        return {arr[k].name.decode(): dict(launches=arr[k].launches, ms=arr[k].total_ms, bytes=arr[k].alg_bytes)
                for k in range(min(n.value, 64))}


def check32(rc):
    if rc != SUQA_OK:
        raise SuqaError(rc, lib().suqa_hip_f32_last_error().decode())
    return rc


class F32Engine:
    """FP32 state with gates fused into width-qubit blocks applied by MFMA."""

    def __init__(self, nq):
        self.ctx = CTX()
        check32(lib().suqa_hip_f32_setup(nq, C.byref(self.ctx)))
        self.nq, self.N = nq, 1 << nq

    def close(self):
        if self.ctx:
            lib().suqa_hip_f32_clear(self.ctx)
            self.ctx = CTX()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def init_state(self):
        check32(lib().suqa_hip_f32_init_state(self.ctx))

    def run_gates(self, gates, width, n=None, start=0):
        n = len(gates) - start if n is None else n
        ptr = C.cast(C.addressof(gates) + start * C.sizeof(Gate), C.POINTER(Gate))
        nb = U32()
        check32(lib().suqa_hip_f32_run_gates(self.ctx, ptr, n, width, C.byref(nb)))
        return nb.value

    def run_paulis(self, ops, width):
        nb = U32()
        check32(lib().suqa_hip_f32_run_paulis(self.ctx, ops, len(ops), width, C.byref(nb)))
        return nb.value

    def qft_inverse(self, qs, width):
        check32(lib().suqa_hip_f32_qft_inverse(self.ctx, _u32(qs), len(qs), width))

    @property
    def passes(self):
        """Application mode: 0 = one block per launch, 1 = block passes,
        2 = block passes with sparse Pauli rotations, 3 = gate passes (no
        fusion: the FP64 engine's sweeps in FP32 tile passes).  True/False
        are 1/0."""
        v = U32()
        check32(lib().suqa_hip_f32_get_passes(self.ctx, C.byref(v)))
        return v.value

    @passes.setter
    def passes(self, mode):
        check32(lib().suqa_hip_f32_set_passes(self.ctx, int(mode)))

    @property
    def tile_bits(self):
        """Tile qubits of a mode-3 gate pass (12 or 13)."""
        v = U32()
        check32(lib().suqa_hip_f32_get_tile_bits(self.ctx, C.byref(v)))
        return v.value

    @tile_bits.setter
    def tile_bits(self, bits):
        check32(lib().suqa_hip_f32_set_tile_bits(self.ctx, int(bits)))

    def autotune(self, gates, n=None):
        """Times the sample in modes 0, 1 at widths 4..6 and in mode 3 at tile
        bits 12, 13; leaves the fastest mode set and returns (best block width,
        {"blocks4": ms per gate, ..., "passes6": ..., "gates12": ..., "gates13": ...})."""
        n = len(gates) if n is None else n
        best = U32()
        t = (C.c_double * 8)()
        check32(lib().suqa_hip_f32_autotune(self.ctx, gates, n, C.byref(best), t))
        out = {f"{m}{4 + k}": t[3 * i + k] for i, m in enumerate(("blocks", "passes")) for k in range(3)}
        out.update({f"gates{12 + k}": t[6 + k] for k in range(2)})
        return best.value, out

    def vnorm(self):
        v = C.c_double()
        check32(lib().suqa_hip_f32_vnorm(self.ctx, C.byref(v)))
        return v.value

    def measure_qbits_deferred(self, qs, rdoubs):
        """Queued measurements; returns the first outcome slot (outcomes: read_outcomes)."""
        rd = (C.c_double * max(1, len(rdoubs)))(*rdoubs)
        slot = U32()
        check(lib().suqa_hip_measure_qbits_deferred(self.ctx, _u32(qs), rd, len(qs), C.byref(slot)))
        return slot.value

    def read_outcomes(self, slot, n):
        out = (U32 * max(1, n))()
        check(lib().suqa_hip_read_outcomes(self.ctx, slot, n, out))
        return [out[k] for k in range(n)]

    def x_if(self, qs, slot):
        check(lib().suqa_hip_apply_x_if(self.ctx, _u32(qs), len(qs), slot, self.gc_mask))

    def qms_apply_W_deferred(self, q_acc, w_mask_enew, bm_enes0, c_factor, eold_slot):
        check(lib().suqa_hip_qms_apply_W_deferred(self.ctx, q_acc, w_mask_enew, bm_enes0, c_factor, eold_slot))

    def prob_filter(self, qs, q_mask):
        p = C.c_double()
        check32(lib().suqa_hip_f32_prob_filter(self.ctx, _u32(qs), _u32(q_mask), len(qs), C.byref(p)))
        return p.value

    def measure(self, q, rdoub):
        c = U32()
        check32(lib().suqa_hip_f32_measure_qbit(self.ctx, q, rdoub, C.byref(c)))
        return c.value

    def fidelity(self, ref):
        f = C.c_double()
        check32(lib().suqa_hip_f32_fidelity(self.ctx, ref.ctx, C.byref(f)))
        return f.value

    def get(self):
        re = np.empty(self.N, dtype=np.float32)
        im = np.empty(self.N, dtype=np.float32)
        check32(lib().suqa_hip_f32_download(self.ctx, re.ctypes.data_as(C.POINTER(C.c_float)),
                                            im.ctypes.data_as(C.POINTER(C.c_float))))
        return re, im

    def synchronize(self):
        check32(lib().suqa_hip_f32_synchronize(self.ctx))

    def timing(self, on):
        check32(lib().suqa_hip_f32_timing_enable(self.ctx, 1 if on else 0))

    def timing_reset(self):
        check32(lib().suqa_hip_f32_timing_reset(self.ctx))

    def timing_report(self):
        """{family: launches, ms, bytes (algorithmic HBM), flops (MFMA)} of the block kernels."""
        arr = (KStat * 64)()
        n = U32()
        check32(lib().suqa_hip_f32_timing_report(self.ctx, arr, 64, C.byref(n)))
        return {arr[k].name.decode(): dict(launches=arr[k].launches, ms=arr[k].total_ms, bytes=arr[k].alg_bytes,
                                           flops=arr[k].flops) for k in range(min(n.value, 64))}


def fuse_plan(gates, nq, width, n=None):
    """Host-only fusion plan: list of (qubits, gates_in, matrix as complex ndarray)."""
    n = len(gates) if n is None else n
    nb = U32()
    check(lib().suqa_hip_fuse_plan(gates, n, nq, width, 0, C.byref(nb), None, None, None))
    B = nb.value
    qs = np.zeros(6 * B, dtype=np.uint32)
    gi = np.zeros(B, dtype=np.uint32)
    D = 1 << width
    mats = np.zeros(2 * B * D * D)
    check(lib().suqa_hip_fuse_plan(gates, n, nq, width, B, C.byref(nb), qs.ctypes.data_as(PU32),
                                   gi.ctypes.data_as(PU32), mats.ctypes.data_as(PD)))
    m = mats.reshape(B, D * D, 2)
    m = (m[..., 0] + 1j * m[..., 1]).reshape(B, D, D)
    return [(list(qs[6 * b:6 * b + width]), int(gi[b]), m[b]) for b in range(B)]


def pauli_ops(seq):
    """ctypes PauliOp array from (qs, types, phase, gc_mask) tuples."""
    arr = (PauliOp * len(seq))()
    for a, (qs, types, phase, gc) in zip(arr, seq):
        if len(qs) != len(types) or len(qs) > 8:
            raise ValueError("a Pauli op needs matching qs/types of length <= 8")
        a.w = len(qs)
        for k, (q, t) in enumerate(zip(qs, types)):
            a.qs[k], a.types[k] = q, t
        a.gc_mask, a.phase = gc, phase
    return arr


def gates_from(seq):
    """Converts a sequence of objects with kind/target/nctrl/ctrl/angle into a Gate array."""
    seq = list(seq)
    arr = (Gate * len(seq))()
    for k, g in enumerate(seq):
        arr[k].kind = g.kind
        arr[k].target = g.target
        arr[k].nctrl = g.nctrl
        for j in range(3):
            arr[k].ctrl[j] = g.ctrl[j]
        arr[k].angle = g.angle
    return arr
