/*
 * suqa_hip.h -- C-ABI of the MI355X (gfx950) state-vector engine.
 *
 * This is the drop-in boundary below the reference's C++ `suqa::` API
 * (include/suqa.cuh:94-176 in the reference).  The C++ mirror in
 * suqa_amd/host/suqa.hpp keeps the reference signatures and calls these
 * entry points; any FFI (ctypes, cgo, JNI) can bind them directly.
 *
 * Conventions
 *   - Every entry point returns an int status (SUQA_OK == 0).  A non-zero
 *     status leaves a message in suqa_hip_last_error().  The C++ mirror turns
 *     it into std::runtime_error, the reference's exception type.
 *   - Plain pointers and sizes only.  Qubit q is bit q of the basis index.
 *   - gc_mask is the reference's global control mask (suqa.cu:36): every
 *     gate taking it acts only on basis states whose gc_mask bits are all 1.
 *   - State layout (the de-facto ABI of complex_defines.cuh:22-45):
 *     one device allocation of 2*N doubles, re = data, im = data + N.
 *   - One HIP stream per context; calls on one context are serialized on it.
 *     Entry points that return a value (vnorm, measure, prob_filter)
 *     synchronize that stream.  Contexts are not thread-safe.
 *   - Reductions are deterministic: a fixed two-level tree for a given nq.
 */
#ifndef SUQA_HIP_H
#define SUQA_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SUQA_OK               0
#define SUQA_ERR_ARG         -1  /* invalid argument (reference: std::runtime_error) */
#define SUQA_ERR_UNSUPPORTED -2  /* Pauli pattern the reference rejects (suqa.cu:555,563,579) */
#define SUQA_ERR_HIP         -3  /* HIP runtime failure */
#define SUQA_ERR_NOMEM       -4  /* device allocation failed */

#define SUQA_PAULI_ID 0u         /* suqa.cuh:60-63 */
#define SUQA_PAULI_X  1u
#define SUQA_PAULI_Y  2u
#define SUQA_PAULI_Z  3u

typedef struct suqa_hip_ctx suqa_hip_ctx;

/* Message of the last failing call on this thread ("" if none). */
const char* suqa_hip_last_error(void);

/* ---- lifecycle ------------------------------------------------------- */
/* suqa::setup (suqa.cu:792-844): allocate 2^nq amplitudes on the current
 * device and set |0...0>.  1 <= nq <= 32. */
int suqa_hip_setup(uint32_t nq, suqa_hip_ctx** out);
/* suqa::clear (suqa.cu:846-866): free everything owned by the context. */
int suqa_hip_clear(suqa_hip_ctx* ctx);
/* suqa::allocate_state (suqa.cu:772-788): resize to 2^nq and set |0...0>. */
int suqa_hip_allocate_state(suqa_hip_ctx* ctx, uint32_t nq);
/* suqa::init_state (suqa.cu:70-76). */
int suqa_hip_init_state(suqa_hip_ctx* ctx);
/* suqa::vnorm (suqa.cu:78-92): sqrt(sum |psi_i|^2). */
int suqa_hip_vnorm(suqa_hip_ctx* ctx, double* norm);
/* suqa::vnormalize (suqa.cu:97-111): multiply by 1./vnorm(). */
int suqa_hip_vnormalize(suqa_hip_ctx* ctx);

/* ---- gates (each one pass over the touched amplitudes) ---------------- */
/* suqa::apply_x(uint) (suqa.cu:115-127) */
int suqa_hip_apply_x(suqa_hip_ctx* ctx, uint32_t q, uint32_t gc_mask);
/* suqa::apply_y(uint) (suqa.cu:136-148) */
int suqa_hip_apply_y(suqa_hip_ctx* ctx, uint32_t q, uint32_t gc_mask);
/* suqa::apply_h(uint) (suqa.cu:207-218) */
int suqa_hip_apply_h(suqa_hip_ctx* ctx, uint32_t q, uint32_t gc_mask);
/* suqa::apply_u1(uint,uint,double) (suqa.cu:271-285); apply_z/s/t/tdg
 * (suqa.cu:156-262) are u1 with pi, pi/2, pi/4, -pi/4.  q_mask selects the
 * |1> (1) or |0> (0) component. */
int suqa_hip_apply_u1(suqa_hip_ctx* ctx, uint32_t q, uint32_t q_mask, double phase,
                      uint32_t gc_mask);
/* suqa::apply_cx / apply_mcx (suqa.cu:289-344): X on target when control k
 * reads ctrl_vals[k] (NULL: all 1). */
int suqa_hip_apply_mcx(suqa_hip_ctx* ctx, const uint32_t* ctrls, const uint32_t* ctrl_vals,
                       uint32_t nctrl, uint32_t target, uint32_t gc_mask);
/* suqa::apply_cu1 / apply_mcu1 (suqa.cu:346-393). */
int suqa_hip_apply_mcu1(suqa_hip_ctx* ctx, const uint32_t* ctrls, const uint32_t* ctrl_vals,
                        uint32_t nctrl, uint32_t target, double phase, uint32_t gc_mask);
/* suqa::apply_swap (suqa.cu:396-416). */
int suqa_hip_apply_swap(suqa_hip_ctx* ctx, uint32_t q1, uint32_t q2, uint32_t gc_mask);
/* suqa::apply_phase_list (suqa.cu:419-454, suqa.cuh:157): phases[k] on the
 * states whose bits q0..q0+q_size-1 read k (documented intent: all gc_mask
 * bits set; the reference GPU kernel instead tests any gc bit).
 * 1 <= q_size <= 7, nphases == 2^q_size. */
int suqa_hip_apply_phase_list(suqa_hip_ctx* ctx, uint32_t q0, uint32_t q_size,
                              const double* phases, uint32_t nphases, uint32_t gc_mask);
/* suqa::apply_pauli_TP_rotation (suqa.cu:467-604): the reference operator,
 * bit for bit, for every pattern the reference accepts (weights 1, 2 and
 * Z(XX|YY|ZZ)); SUQA_ERR_UNSUPPORTED where the reference throws. */
int suqa_hip_apply_pauli_TP_rotation(suqa_hip_ctx* ctx, const uint32_t* qs, const uint32_t* types,
                                     uint32_t n, double phase, uint32_t gc_mask);
/* Extension (the reference's TODO at suqa.cu:459): exp(i*phase*P) for any
 * Pauli string, P = prod_k sigma_{types[k]} on qs[k], in one pass. */
int suqa_hip_apply_pauli_rotation_general(suqa_hip_ctx* ctx, const uint32_t* qs,
                                          const uint32_t* types, uint32_t n, double phase,
                                          uint32_t gc_mask);
/* suqa::apply_qft / apply_qft_inverse (suqa.cu:635-661): no final swaps; the
 * crm phase table is filled once per context by the first call
 * (suqa.cu:621-625 does it once per process). */
int suqa_hip_apply_qft(suqa_hip_ctx* ctx, const uint32_t* qs, uint32_t n, uint32_t gc_mask);
int suqa_hip_apply_qft_inverse(suqa_hip_ctx* ctx, const uint32_t* qs, uint32_t n,
                               uint32_t gc_mask);

/* ---- measurement ------------------------------------------------------ */
/* suqa::measure_qbit (suqa.cu:673-696): c = (rdoub < P(q=1)); collapse and
 * renormalize.  One fused reduction + one fused collapse/scale pass. */
int suqa_hip_measure_qbit(suqa_hip_ctx* ctx, uint32_t q, double rdoub, uint32_t* c);
/* suqa::apply_reset(uint,double) (suqa.cu:730-738): measure, X if 1.  On a
 * state of at most 2^14 amplitudes the measurement, collapse and X are one
 * launch and the outcome stays on the device (nothing waits for it). */
int suqa_hip_apply_reset(suqa_hip_ctx* ctx, uint32_t q, double rdoub, uint32_t gc_mask);
/* suqa::measure_qbits (suqa.cu:699-702): measure_qbit on qs[k] with rdoubs[k],
 * in order, outcomes into cs[0..n).  On a state of at most 2^14 amplitudes
 * the n measurements are one launch and one wait; with cs == NULL (outcomes
 * not wanted) they are one launch and nothing waits.  On a state of at most
 * 2^10 amplitudes that is one pass tile they join the queued gates instead:
 * the queue and the measurements are one launch. */
int suqa_hip_measure_qbits(suqa_hip_ctx* ctx, const uint32_t* qs, const double* rdoubs, uint32_t n, uint32_t* cs);
/* suqa::apply_reset(bmReg, vector<double>) (suqa.cu:752-757): apply_reset on
 * qs[k] with rdoubs[k], in order; one launch and no wait on a state of at
 * most 2^14 amplitudes. */
int suqa_hip_apply_resets(suqa_hip_ctx* ctx, const uint32_t* qs, const double* rdoubs, uint32_t n, uint32_t gc_mask);
/* Extension (no counterpart in the reference): deferred measurements, for
 * callers that need an outcome on the device before they need it on the host
 * (qms::'s energy readout, which the X flips and apply_W read).  measure_qbits'
 * measurements (draws rdoubs[k]) with their outcomes in outcome slots
 * *slot .. *slot + n - 1 of a ring of SUQA_HIP_OUTCOME_SLOTS per context, and
 * nothing waits: on a state of at most 2^10 amplitudes that is one pass tile
 * they are queued beside the gates, on other states of at most 2^14 amplitudes
 * they are launched at once.  A slot keeps its outcome until that many later
 * slots have been handed out.  SUQA_ERR_UNSUPPORTED on larger states. */
#define SUQA_HIP_OUTCOME_SLOTS 1024u
int suqa_hip_measure_qbits_deferred(suqa_hip_ctx* ctx, const uint32_t* qs, const double* rdoubs, uint32_t n,
                                    uint32_t* slot);
/* Outcomes of slots slot, slot + 1, ... (n of them, ring order) into
 * out[0..n): applies what is queued and waits for it. */
int suqa_hip_read_outcomes(suqa_hip_ctx* ctx, uint32_t slot, uint32_t n, uint32_t* out);
/* X on qs[k] over the indices with the gc bits set, where outcome slot + k
 * reads 1 (the flips of a measured register, qms.cuh:404-406): queued on a
 * state of one small pass tile, else one launch (states of at most 2^14
 * amplitudes). */
int suqa_hip_apply_x_if(suqa_hip_ctx* ctx, const uint32_t* qs, uint32_t n, uint32_t slot, uint32_t gc_mask);
/* For k = 0 .. n-1 in order: apply_reset(qs[k], rdoubs[k]) over the gc bits,
 * then X on xqs[k] over the gc bits where outcome slot + k reads 1 -- the
 * reference's interleaving of the E_old register's resets with the flips of
 * the measured E_new register (qms.cuh:431-434, the two-register layout).
 * The same operations as those calls in that order (bit-identical); on a
 * state of at most 2^14 amplitudes that is not one pass tile, one launch
 * instead of 2n. */
int suqa_hip_resets_then_x_if(suqa_hip_ctx* ctx, const uint32_t* qs, const double* rdoubs, const uint32_t* xqs,
                              uint32_t n, uint32_t slot, uint32_t gc_mask);
/* suqa::prob_filter (suqa.cu:705-727): sum |psi_i|^2 over i whose bits qs[k]
 * read q_mask[k]. */
int suqa_hip_prob_filter(suqa_hip_ctx* ctx, const uint32_t* qs, const uint32_t* q_mask,
                         uint32_t n, double* prob);

/* No counterpart in the reference: suqa::prob_filter (suqa.cu:705-727) on a
 * state the caller holds, for callers that cannot allocate device memory
 * (code under HIP graph capture, serving loops, harnesses that forbid
 * hipMalloc).  It allocates nothing: the reduction partials live in a fixed
 * device array of the library (4 MiB per GPU, held until the process ends),
 * so 1 <= nq <= 26.  d_re, d_im: the state's 2^nq real and imaginary parts in
 * device memory; d_prob: one device double that receives the probability.
 * Asynchronous on `stream` (a hipStream_t; NULL: the null stream), no host
 * synchronization; the same kernels as suqa_hip_prob_filter, so the same
 * result bit for bit.  Calls that overlap on one GPU share the array and
 * race: serialize them (one stream, or events). */
int suqa_hip_prob_filter_noalloc(const double* d_re, const double* d_im, uint32_t nq, const uint32_t* qs,
                                 const uint32_t* q_mask, uint32_t n, double* d_prob, void* stream);

/* ---- QMS -------------------------------------------------------------- */
/* qms::apply_W (qms.cuh:220-329): acceptance rotation on ancilla q_acc (the
 * most significant qubit) driven by the energy register read by
 * (i & W_mask_Enew) >> bm_enes0.  fs1 = exp(-dE/2), fs2 = sqrt(1-fs1^2) are
 * evaluated on the host with libm (as the reference CPU path does). */
int suqa_hip_qms_apply_W(suqa_hip_ctx* ctx, uint32_t q_acc, uint32_t W_mask_Enew,
                         uint32_t bm_enes0, double c_factor, uint32_t curr_E_old);
/* Extension: the same with E_old read on the device from the outcome slots
 * eold_slot .. eold_slot + r - 1 (LSB first; r = the register's width) of
 * deferred measurements (suqa_hip_measure_qbits_deferred). */
int suqa_hip_qms_apply_W_deferred(suqa_hip_ctx* ctx, uint32_t q_acc, uint32_t W_mask_Enew, uint32_t bm_enes0,
                                  double c_factor, uint32_t eold_slot);
/* The same rotation in the two-register layout (the commented bm_enes_old
 * code of qms.cuh:72-85, 134-136, 201-202; extension): E_old is read from
 * the register (W_mask_Eold, offset bm_enes_old0) of each basis state instead
 * of a classical value.  Registers: contiguous, equal width <= 8 qubits,
 * disjoint, below the acceptance qubit (the most significant). */
int suqa_hip_qms_apply_W2(suqa_hip_ctx* ctx, uint32_t q_acc, uint32_t W_mask_Enew, uint32_t bm_enes_new0,
                          uint32_t W_mask_Eold, uint32_t bm_enes_old0, double c_factor);

/* ---- state access (ComplexVec view, suqa.cu:17) ----------------------- */
int suqa_hip_state_info(suqa_hip_ctx* ctx, uint32_t* nq, uint64_t* N, double** dev_re,
                        double** dev_im);
int suqa_hip_download(suqa_hip_ctx* ctx, double* re, double* im);
int suqa_hip_upload(suqa_hip_ctx* ctx, const double* re, const double* im);
int suqa_hip_synchronize(suqa_hip_ctx* ctx);
/* Extension (no counterpart in the reference): a checksum of the state,
 * out3 = {sum re, sum im, sum_i i |psi_i|^2 / N}, FP64 tree sums over the
 * whole state (deterministic for a given nq). */
int suqa_hip_checksum(suqa_hip_ctx* ctx, double* out3);

/* ---- captured launch sequences (extension) ----------------------------
 * No counterpart in the reference, which launches every kernel from the
 * host.  graph_begin applies the queued gates, then captures the context
 * stream: gate calls until graph_end are recorded (the passes they form),
 * not run; graph_end returns an id; graph_launch applies the queued gates,
 * then replays the sequence with the same results as the calls themselves.
 * Entry points that read a value back (norms, probabilities, measurements)
 * fail while a capture is open.  A graph holds the state's device
 * addresses: allocate_state drops every graph of the context. */
int suqa_hip_graph_begin(suqa_hip_ctx* ctx);
int suqa_hip_graph_end(suqa_hip_ctx* ctx, uint32_t* graph);
int suqa_hip_graph_launch(suqa_hip_ctx* ctx, uint32_t graph);

/* ---- workload runner and timing (extensions) -------------------------- */
/* Gate kinds of the random-circuit workload. RX(a) = exp(-i a X/2),
 * RZ(a) = exp(-i a Z/2), U1(a) = diag(1, e^{ia}). */
#define SUQA_GATE_H   0
#define SUQA_GATE_X   1
#define SUQA_GATE_RX  2
#define SUQA_GATE_RZ  3
#define SUQA_GATE_U1  4
#define SUQA_GATE_CX  5
#define SUQA_GATE_MCX 6
typedef struct {
    int32_t  kind;
    uint32_t target;
    uint32_t nctrl;
    uint32_t ctrl[3];
    double   angle;
} suqa_hip_gate;

/* Applies gates[0..ngates) in order, asynchronously on the context stream. */
int suqa_hip_run_gates(suqa_hip_ctx* ctx, const suqa_hip_gate* gates, uint32_t ngates);

/* Gate passes (extension; the reference runs one sweep per gate).  With
 * bits = t in 1..12, gate calls are queued and applied in passes: runs of
 * consecutive gates whose flipped qubits fit in one t-qubit tile are applied
 * to LDS-resident tiles of 2^t amplitudes in one HBM round trip, in order and
 * with the one-sweep arithmetic (bit-identical results).  Every call that
 * reads or exposes the state (norms, probabilities, measurements, apply_W,
 * phase_list, download/upload, state_info, synchronize, timing_report)
 * applies the queue first.  bits = 0 gives one sweep per gate.  A new
 * context takes SUQA_HIP_TILE_BITS from the environment, else 11. */
int suqa_hip_set_tile_bits(suqa_hip_ctx* ctx, uint32_t bits);
int suqa_hip_get_tile_bits(suqa_hip_ctx* ctx, uint32_t* bits);

/* Gate order of the passes (extension; the reference applies gates in call
 * order).  packed = 0 (the default): queued gates run in the order they were
 * queued, bit-identical to one sweep per gate.  packed = 1: each flush packs
 * its passes greedily, a gate moving ahead of earlier queued gates only when
 * it commutes with each of them (neither flips a qubit the other reads): the
 * product is unchanged, the rounding is not (amplitudes within 1e-12 of call
 * order at 30 qubits and 200 layers), and there are fewer passes.  Gates
 * queued before the call keep their order.  The suqa:: C++ API never sets it. */
int suqa_hip_set_gate_order(suqa_hip_ctx* ctx, int packed);
/* Autotuner: applies the gate sample once per candidate tile size (so the
 * state changes: re-initialise it afterwards), times it, keeps the fastest
 * in *best and sets it; ms_per_gate[k] (may be NULL) gets each candidate's
 * wall time per gate. */
int suqa_hip_autotune_tile_bits(suqa_hip_ctx* ctx, const suqa_hip_gate* gates, uint32_t ngates,
                                const uint32_t* candidates, uint32_t ncand, uint32_t* best, double* ms_per_gate);
/* Counters since setup: passes launched, gates applied in passes, gates
 * applied as lone sweeps while passes were on. */
int suqa_hip_pass_stats(suqa_hip_ctx* ctx, uint64_t* passes, uint64_t* pass_gates, uint64_t* single_sweeps);

/* One Pauli-string rotation exp(i phase P), P = prod_k sigma_{types[k]} on
 * qs[k] (k < w <= 8), controlled on the gc_mask qubits: the operation
 * suqa_hip_apply_pauli_rotation_general performs, as a list element. */
typedef struct {
    uint32_t w;
    uint32_t qs[8];
    uint32_t types[8];
    uint32_t gc_mask;
    double   phase;
} suqa_hip_pauli_op;

/* Applies ops[0..nops) in order, one sweep each (a Trotter step list). */
int suqa_hip_run_paulis(suqa_hip_ctx* ctx, const suqa_hip_pauli_op* ops, uint32_t nops);

typedef struct {
    char     name[32];     /* kernel family */
    uint64_t launches;
    double   total_ms;     /* sum of per-launch HIP-event durations */
    double   alg_bytes;    /* sum of algorithmic HBM bytes */
    double   flops;        /* sum of MFMA flops (FP32 block kernels; 0 elsewhere) */
} suqa_hip_kstat;

/* Per-launch HIP events around every kernel launch on the context stream. */
int suqa_hip_timing_enable(suqa_hip_ctx* ctx, int on);
int suqa_hip_timing_reset(suqa_hip_ctx* ctx);
/* Synchronizes, then fills up to max entries; *n receives the count. */
int suqa_hip_timing_report(suqa_hip_ctx* ctx, suqa_hip_kstat* out, uint32_t max, uint32_t* n);

/* Device copy dst <- src of nbytes on the context stream (HBM calibration). */
int suqa_hip_copy_kernel(suqa_hip_ctx* ctx, void* dst, const void* src, uint64_t nbytes);
/* Device buffer helpers for the calibration above. */
int suqa_hip_device_alloc(uint64_t nbytes, void** ptr);
int suqa_hip_device_free(void* ptr);

/* ---- FP32 state with fused blocks (the precision pattern; extension) --
 * No counterpart in the reference.  Gates of a list are fused greedily, with
 * commutation lookahead, into dense blocks of exactly `width` qubits
 * (4..6; built in FP64, rounded to FP32) and applied with FP32 MFMA.  Norms,
 * probabilities, the collapse scale and fidelities accumulate in FP64.
 * Requires 9 <= nq <= 33 and nq >= width + 5. */
typedef struct suqa_hip_f32 suqa_hip_f32;

const char* suqa_hip_f32_last_error(void);
int suqa_hip_f32_setup(uint32_t nq, suqa_hip_f32** out);
int suqa_hip_f32_clear(suqa_hip_f32* c);
int suqa_hip_f32_init_state(suqa_hip_f32* c);
/* Application mode of the fused blocks: 0 (the default) fuses the whole list
 * and applies each block in its own launch (one HBM round trip per block);
 * 1 cuts the list into passes of consecutive ops whose targets fit a 12-qubit
 * tile, fuses each pass and applies all of its blocks in one round trip;
 * 2 is mode 1 with the Pauli rotations of suqa_hip_f32_run_paulis applied as
 * pair updates in LDS instead of fused MFMA blocks (a rotation has two
 * nonzeros per row); 3 fuses nothing: every gate becomes the FP64 engine's
 * sweep (the gates of suqa_hip_run_gates, the rotations of
 * suqa_hip_run_paulis, the H and phase gates of suqa_hip_apply_qft_inverse)
 * and the gate-pass scheduler of the FP64 engine applies them in FP32 VALU
 * arithmetic, in passes over LDS tiles of suqa_hip_f32_set_tile_bits qubits
 * (the width argument is then ignored, *nblocks receives the number of
 * passes, and Pauli strings keep the FP64 engine's limits: weight <= 8, no
 * control on a qubit of the string).  Results agree to FP32 rounding (the
 * arithmetic differs). */
int suqa_hip_f32_set_passes(suqa_hip_f32* c, uint32_t on);
int suqa_hip_f32_get_passes(const suqa_hip_f32* c, uint32_t* on);
/* Mode 3's pass tile: 12 (the default; 32 KiB of LDS per workgroup) or 13
 * qubits (64 KiB). */
int suqa_hip_f32_set_tile_bits(suqa_hip_f32* c, uint32_t bits);
int suqa_hip_f32_get_tile_bits(const suqa_hip_f32* c, uint32_t* bits);
/* Fuses gates[0..ngates) into width-qubit blocks and applies them in order;
 * *nblocks (may be NULL) receives the number of blocks. */
int suqa_hip_f32_run_gates(suqa_hip_f32* c, const suqa_hip_gate* gates, uint32_t ngates, uint32_t width,
                           uint32_t* nblocks);
/* The same fused application for Pauli-string rotations: each string's
 * non-identity qubits (at most `width`) carry the block matrix, and the
 * gc_mask controls select the tiles it is applied to (ops sharing a control
 * mask fuse into one controlled block; tiles with a control at 0 are not
 * read).  Controls must not be qubits of the string, nor qubits 0 or 1. */
int suqa_hip_f32_run_paulis(suqa_hip_f32* c, const suqa_hip_pauli_op* ops, uint32_t nops, uint32_t width,
                            uint32_t* nblocks);
/* suqa::apply_qft_inverse (suqa.cu:649-661, no final swaps) on qs[0..n),
 * its H and controlled-phase gates fused into width-qubit blocks. */
int suqa_hip_f32_qft_inverse(suqa_hip_f32* c, const uint32_t* qs, uint32_t n, uint32_t width);
/* Times the gate sample on this context's state (which it changes) in modes
 * 0 and 1 at every valid width and in mode 3 at tile bits 12 and 13, leaves
 * the fastest mode (and mode 3's tile bits) set and returns the fastest
 * block width of modes 0/1 in *best_width; ms_per_gate (8 entries) gets the
 * wall time per gate: [3 * mode + width - 4] for modes 0, 1 and
 * [6 + bits - 12] for mode 3 (-1 where it does not fit). */
int suqa_hip_f32_autotune(suqa_hip_f32* c, const suqa_hip_gate* gates, uint32_t ngates, uint32_t* best_width,
                          double* ms_per_gate);
int suqa_hip_f32_vnorm(suqa_hip_f32* c, double* norm);
int suqa_hip_f32_prob_filter(suqa_hip_f32* c, const uint32_t* qs, const uint32_t* q_mask, uint32_t n, double* prob);
/* measure_qbit semantics (suqa.cu:673-696) on the FP32 state, FP64 probabilities. */
int suqa_hip_f32_measure_qbit(suqa_hip_f32* c, uint32_t q, double rdoub, uint32_t* cbit);
/* |<ref|psi>|^2 / (<ref|ref> <psi|psi>) against an FP64 engine state of the
 * same size (FP64 sums). */
int suqa_hip_f32_fidelity(suqa_hip_f32* c, suqa_hip_ctx* ref, double* fidelity);
int suqa_hip_f32_download(suqa_hip_f32* c, float* re, float* im);
int suqa_hip_f32_synchronize(suqa_hip_f32* c);
/* Per-launch HIP events around every block kernel (k_block<K> as "blockK",
 * k_bpass<K> as "bpassK", mode 3's k_tile<float> as "tile_pass32"):
 * algorithmic HBM bytes (16 B per amplitude of the tiles a launch visits:
 * FP32 re and im read and written) and MFMA flops (8 * 2^K per amplitude per
 * dense block; 0 for the gate passes). */
int suqa_hip_f32_timing_enable(suqa_hip_f32* c, int on);
int suqa_hip_f32_timing_reset(suqa_hip_f32* c);
int suqa_hip_f32_timing_report(suqa_hip_f32* c, suqa_hip_kstat* out, uint32_t max, uint32_t* n);
/* Host only (no GPU): the fusion plan.  For each block b < *nblocks (at most
 * max_blocks): qubits[6b..6b+width) sorted, gates_in[b], and the FP64 matrix
 * mats[b * 2 * 4^width ...] as interleaved (re, im), row-major over the
 * block index sum_j bit(qubits[j]) 2^j. */
int suqa_hip_fuse_plan(const suqa_hip_gate* gates, uint32_t ngates, uint32_t nq, uint32_t width,
                       uint32_t max_blocks, uint32_t* nblocks, uint32_t* qubits, uint32_t* gates_in, double* mats);

#ifdef __cplusplus
}
#endif
#endif
