"""Synthetic workloads of the benchmark, generated with the reference's RNG.

PCG32 follows pcg32.h:53-68 / 121-130 (O'Neill's PCG-XSH-RR, Jakob's tiny
version) and the `pcg` wrapper of Rand.hpp:249-291 (seed with stream 1,
doub() fills 32 mantissa bits, randint(a,b) = a + next() % (b-a)).

random_circuit(n, layers, seed): the gate-throughput workload.
  |0>, H on every qubit, then per layer n gates:
    kind = randint(0,7) over {H, X, RX, RZ, U1, CX, MCX(3 controls)}
    target = randint(0,n); controls drawn with randint(0,n), distinct from
    the target and each other; angle = 2*pi*doub() for RX, RZ, U1.
  then min(8,n) distinct measured qubits, one doub() each.
trotter_terms(nsys, nterms, seed): Pauli strings of weight randint(2,5),
  distinct qubits in [0,nsys), types randint(1,4), coefficient 2*doub()-1.
"""
import math

MASK64 = (1 << 64) - 1
PCG_MULT = 0x5851F42D4C957F2D


class Pcg32:
    def __init__(self, seed, seq=1):
        self.state = 0
        self.inc = ((seq << 1) | 1) & MASK64
        self.next()
        self.state = (self.state + seed) & MASK64
        self.next()

    def next(self):
        old = self.state
        self.state = (old * PCG_MULT + self.inc) & MASK64
        xs = (((old >> 18) ^ old) >> 27) & 0xFFFFFFFF
        rot = old >> 59
        return ((xs >> rot) | (xs << ((-rot) & 31))) & 0xFFFFFFFF

    def doub(self):
        # (u32 << 20 | exponent of 1.0) - 1.0: exactly u32 * 2^-32
        return self.next() * 2.0 ** -32

    def randint(self, a, b):
        return a + self.next() % (b - a)


H, X, RX, RZ, U1, CX, MCX = range(7)
GATE_NAMES = ["H", "X", "RX", "RZ", "U1", "CX", "MCX"]


class G:
    __slots__ = ("kind", "target", "nctrl", "ctrl", "angle")

    def __init__(self, kind, target, ctrl, angle):
        self.kind, self.target, self.nctrl = kind, target, len(ctrl)
        self.ctrl = list(ctrl) + [0] * (3 - len(ctrl))
        self.angle = angle


def random_circuit(n, layers, seed):
    r = Pcg32(seed)
    gates = []
    for _ in range(layers):
        for _ in range(n):
            kind = r.randint(0, 7)
            t = r.randint(0, n)
            nc = 1 if kind == CX else (3 if kind == MCX else 0)
            ctrl = []
            while len(ctrl) < nc:
                c = r.randint(0, n)
                if c != t and c not in ctrl:
                    ctrl.append(c)
            ang = 2.0 * math.pi * r.doub() if kind in (RX, RZ, U1) else 0.0
            gates.append(G(kind, t, ctrl, ang))
    mq = []
    while len(mq) < min(8, n):
        q = r.randint(0, n)
        if q not in mq:
            mq.append(q)
    mr = [r.doub() for _ in mq]
    return gates, mq, mr


def gate_bytes(g, n):
    """Algorithmic HBM bytes of one plain FP64 gate: 32 B per touched amplitude."""
    N = 1 << n
    if g.kind in (H, X, RX, RZ):
        return 32 * N
    if g.kind in (U1, CX):
        return 16 * N
    return 4 * N   # MCX with 3 controls: N/8 amplitudes


MEASURE_BYTES_PER_AMP = 32   # read the bit-1 half, write the zero half, scale the kept half


class Term:
    __slots__ = ("w", "qs", "types", "coef")


def trotter_terms(nsys, nterms, seed):
    r = Pcg32(seed)
    out = []
    for _ in range(nterms):
        t = Term()
        t.w = r.randint(2, 5)
        qs = []
        while len(qs) < t.w:
            q = r.randint(0, nsys)
            if q not in qs:
                qs.append(q)
        t.qs = qs
        t.types = [r.randint(1, 4) for _ in range(t.w)]
        t.coef = 2.0 * r.doub() - 1.0
        out.append(t)
    return out


def trotter_step_ops(terms, dt, gc_mask=0):
    """One first-order Trotter step exp(-i dt H), H = sum_t coef_t P_t, as
    Pauli rotations exp(i (-coef dt) P) controlled on gc_mask, in term order."""
    return [(list(t.qs), list(t.types), -t.coef * dt, gc_mask) for t in terms]


def qpe_trotter(eng, nsys, nreg, terms, dt, steps, width=None, graphs=True):
    """Phase estimation of U = (Trotter step)^steps on an nsys-qubit system
    with an nreg-qubit register on qubits nsys..nsys+nreg-1: H on the
    register; for trg = nreg-1 .. 0, U^(2^(nreg-1-trg)) controlled on
    register qubit trg; the reference's no-swap inverse QFT on the register
    (suqa.cu:649-661).  eng is a capi.Engine (FP64, one sweep per rotation)
    or a capi.F32Engine (width-qubit fused blocks).  Returns the 2^nreg-bin
    histogram from prob_filter in FP64, bin k = register value k (LSB on the
    first register qubit).  FP64 with graphs: each register qubit's controlled
    Trotter step is captured once (suqa_hip_graph_*) and the capture is
    replayed 2^(nreg-1-trg) * steps times (the same launches, the same
    results)."""
    from suqa_amd import capi
    reg = [nsys + k for k in range(nreg)]
    f32 = isinstance(eng, capi.F32Engine)
    had = capi.gates_from([G(H, q, [], 0.0) for q in reg])
    if f32:
        eng.run_gates(had, width)
    else:
        eng.run_gates(had)
    for trg in range(nreg - 1, -1, -1):
        powr = 1 << (nreg - 1 - trg)
        ops = capi.pauli_ops(trotter_step_ops(terms, dt, 1 << reg[trg]))
        if f32:
            allops = capi.pauli_ops(trotter_step_ops(terms, dt, 1 << reg[trg]) * (powr * steps))
            eng.run_paulis(allops, width)
        elif graphs:
            eng.graph_begin()
            eng.run_paulis(ops)
            g = eng.graph_end()
            for _ in range(powr * steps):
                eng.graph_launch(g)
        else:
            for _ in range(powr * steps):
                eng.run_paulis(ops)
    if f32:
        eng.qft_inverse(reg, width)
    else:
        eng.qft_inverse(reg)
    return [eng.prob_filter(reg, [(k >> j) & 1 for j in range(nreg)]) for k in range(1 << nreg)]

