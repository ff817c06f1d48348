/*
 * dqsim_gpu.h -- C ABI of the MI355X (gfx950) statevector engine.
 *
 * These entry points replace the Rayon kernels of the reference's
 * src/engine.rs one for one; a Rust/PyO3 host binds them through an
 * `extern "C"` block.  Plain pointers and sizes only; no torch or HIP types
 * cross the boundary.
 *
 * State layout in HBM: one contiguous array of 2^n amplitudes, interleaved
 * (re, im), amplitude index bit q <-> qubit q (qubit 0 = LSB), exactly the
 * reference's Vec<Complex64> layout (src/engine.rs:34,113).  Multi-qubit
 * matrices are row-major 2^k x 2^k, (re, im) interleaved doubles, with
 * qubits[0] the most significant bit of the row/column index
 * (src/engine.rs:104-106,116-125).
 *
 * Execution model: gate applications are queued on the state's stream and
 * may be fused (several gates applied in one HBM sweep, in program order, with
 * each gate's arithmetic unchanged).  Every call that returns data to the host
 * (dq_prob_one, dq_marginal_probs, dq_inner, dq_copy_to_host)
 * first flushes the queue and synchronises.  A handle is single-stream and not
 * re-entrant: calls on one handle must not overlap.
 *
 * Return codes: 0 ok, < 0 error; dq_last_error() gives a thread-local message.
 */
#ifndef DQSIM_GPU_H
#define DQSIM_GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct dq_state dq_state;

typedef enum { DQ_C128 = 0, DQ_C64 = 1 } dq_dtype;

enum {
    DQ_OK = 0,
    DQ_ERR_ARG = -1,     /* bad qubit / arity / size: host raises ValueError */
    DQ_ERR_OOM = -2,     /* device allocation failed: host raises MemoryError */
    DQ_ERR_DEVICE = -3,  /* HIP runtime error / no device: RuntimeError */
};

/* Allocates 2^n amplitudes on the current device and sets |0...0>.
 * Replaces `vec![C::new(0.0,0.0); 1<<n]; state[0]=1`
 * (src/monolithic/statevector/simulator.rs:338-339). */
int dq_state_create(int n_qubits, int dtype, dq_state **out);
void dq_state_destroy(dq_state *s);
int dq_state_num_qubits(const dq_state *s);
/* Overwrites the whole state from host memory (2^n (re,im) doubles). */
int dq_state_set(dq_state *s, const double *src);
/* Resets the state to |0...0>. */
int dq_state_reset(dq_state *s);

/* 2x2 gate u (row-major, (re,im)) on `target`, optionally restricted to the
 * subspace where every ctrl_wires[c] has value ctrl_vals[c].
 * Replaces engine::apply_one_qubit (src/engine.rs:24-101). */
int dq_apply_1q(dq_state *s, const double u[8], int target, const int *ctrl_wires,
                const unsigned char *ctrl_vals, int n_ctrl);

/* Dense 2^k x 2^k gate on qubits[0..k) (qubits[0] = MSB), 1 <= k <= 5.
 * Replaces engine::apply_n_qubit (src/engine.rs:109-181): each output
 * amplitude is acc = 0; acc += U[r][c] * v[c] for c ascending. */
int dq_apply_nq(dq_state *s, const double *u, const int *qubits, int k);

/* p1 = sum over amplitudes with bit `qubit` set of |psi_i|^2 (FP64).
 * Synchronous.  First half of engine::measure_qubit (src/engine.rs:197-208);
 * the random draw stays on the host. */
int dq_prob_one(dq_state *s, int qubit, double *p1);

/* Zeroes the half with bit `qubit` != outcome and divides the kept half by
 * `norm` (IEEE division).  Second half of engine::measure_qubit
 * (src/engine.rs:212-254); the host computes norm = max(sqrt(p1),1e-15) or
 * max(sqrt(max(1-p1,0)),1e-15). */
int dq_collapse(dq_state *s, int qubit, int outcome, double norm);

/* out[j] = sum_{i : bits(i, qubits) == j} |psi_i|^2, qubits[0] = MSB of j,
 * 0 <= k <= n, out has 2^k doubles (host).  Replaces engine::marginal_probs
 * (src/engine.rs:368-390) behind SimulationResult.probabilities/counts. */
int dq_marginal_probs(dq_state *s, const int *qubits, int k, double *out);

/* dq_marginal_probs without any device allocation: no hipMalloc, hipFree
 * or allocating library call.  No counterpart in the reference; it is for
 * callers that cannot allocate (code under graph capture, tight serving
 * loops, harnesses that forbid it).  Its work space is a fixed array of
 * 4 MiB + 4 KiB in the library's device memory, one per GPU, held until the
 * process ends: a call needs 512 + 8 * 2^k bytes, plus as much again for
 * partial sums when a bin adds up more than 64 amplitudes, so k <= 18 always
 * fits and k = 19 fits when n - k <= 6 (DQ_ERR_ARG beyond).  The state must have
 * no queued gates (DQ_ERR_ARG otherwise: call dq_synchronize first).  Calls
 * on one GPU share the array and are serialised by a lock; `out` (host, 2^k
 * doubles) is written before the call returns.  Same values as
 * dq_marginal_probs. */
int dq_marginal_probs_noalloc(dq_state *s, const int *qubits, int k, double *out);

/* sum_i conj(psi_i) * other_i with `other` a host array of 2^n (re,im)
 * doubles.  Replaces the fold in SimulationResult.fidelity
 * (src/monolithic/statevector/simulator.rs:184-191). */
int dq_inner(dq_state *s, const double *other, double *re, double *im);

/* Copies the state to host memory (2^n (re,im) doubles; for DQ_C64 states the
 * values are widened).  Backs SimulationResult.statevector
 * (src/monolithic/statevector/simulator.rs:122-125). */
int dq_copy_to_host(dq_state *s, double *dst);

/* ---- observables (no reference counterpart: the adjoint-gradient form) ----
 * A Pauli string is (xmask, zmask): X on xmask's bits, Z on zmask's bits, Y
 * where both are set (Y = i X Z).  coeff holds n_terms (re, im) pairs. */

/* dst = src (amplitudes; both states must have the same size and dtype). */
int dq_state_copy(dq_state *dst, dq_state *src);

/* dst = a (x) b with the qubits interleaved: local qubit k of a becomes
 * qubit a_pos[k] of dst, local qubit k of b qubit b_pos[k]; together they
 * name every qubit of dst (dst->n = a->n + b->n) once.  Each amplitude is
 * one complex product, as merge_blocks (src/distributed/pblock/model.rs:30-58)
 * computes it.  dst must be distinct from a and b; all three share a dtype. */
int dq_state_tensor(dq_state *dst, dq_state *a, const int *a_pos, dq_state *b, const int *b_pos);

/* dst = sum_k coeff_k P_k src  (dst and src distinct). */
int dq_apply_pauli_sum(dq_state *dst, dq_state *src, int n_terms, const uint64_t *xmask,
                       const uint64_t *zmask, const double *coeff);

/* (re, im) = sum_k coeff_k <bra| P_k |ket>, FP64 accumulation; synchronous.
 * With one identity term (0, 0, 1) this is <bra|ket>. */
int dq_pauli_inner(dq_state *bra, dq_state *ket, int n_terms, const uint64_t *xmask,
                   const uint64_t *zmask, const double *coeff, double *re, double *im);

/* Reverse sweep of the adjoint-gradient method (no reference counterpart):
 * for l = n_ops-1 down to 0, grad[l] = Im <lam| G_l |psi> when n_terms[l] > 0
 * (else 0), then psi <- U_l psi and lam <- U_l lam.  Op l acts on
 * qubits[5*l .. 5*l+k[l]) (first = MSB) with the 2^k x 2^k matrix U_l taken
 * from `mats` (row-major complex128, ops concatenated) -- normally the
 * inverse of forward gate l; G_l = sum of n_terms[l] Pauli terms (xmask,
 * zmask, (re, im) coeff, as dq_pauli_inner takes them, ops concatenated)
 * acting on op l's qubits only.  Runs as tile passes that read and write each
 * state once per pass; n >= 8. */
int dq_adjoint_sweep(dq_state *psi, dq_state *lam, int n_ops, const int *k, const int *qubits,
                     const double *mats, const int *n_terms, const uint64_t *xmask, const uint64_t *zmask,
                     const double *coeff, double *grad);

/* Lazy results (no reference counterpart): the k (1..2^20) most probable
 * basis states, p = |amp|^2 as probabilities() computes it, sorted by
 * decreasing p, ties by increasing index; the distribution is never copied
 * to the host (radix select on the device). */
int dq_top_k(dq_state *s, uint64_t k, uint64_t *idx_out, double *p_out);

/* Compares two states of the same size and any dtypes (the complex64 error
 * report): max_i |a_i - b_i|, <a|b>, and the total-variation distance
 * 0.5 * sum_i | |a_i|^2 - |b_i|^2 |, all accumulated in FP64. */
int dq_compare_states(dq_state *a, dq_state *b, double *max_abs, double *inner_re,
                      double *inner_im, double *tvd);

/* Flushes queued gates and waits for the device. */
int dq_synchronize(dq_state *s);

/* Execution options (no reference counterpart; tuning only; 0 keeps a value).
 *   fuse = 0: every gate is its own HBM sweep;
 *   fuse = 1: gates are grouped into LDS-tiled sweeps in strict program
 *             order (bit-identical to the reference's arithmetic);
 *   fuse = 2: as 1, but a sweep may also take later gates on qubits disjoint
 *             from every gate it skipped (default; identical in exact
 *             arithmetic, last-bit rounding may differ);
 *   fuse = 3: complex64 states only: as 2, and runs of gates on at most 5
 *             qubits are fused (FP64 product on the host) into 32x32 blocks
 *             applied as GEMMs on FP32 MFMA.
 *   tile_bits: qubits per tile (6..13);
 *   low_bits: qubits 0..low_bits-1 are in every tile, so each tile is made of
 *             contiguous runs of 2^low_bits amplitudes (1..8). */
int dq_set_options(dq_state *s, int fuse, int tile_bits, int low_bits);

/* Sweep-kernel timing (no reference counterpart; measurement only).  With
 * timing on, every state-sweep kernel launch is bracketed by HIP events on
 * the state's stream; dq_get_timing flushes, synchronises and returns the
 * summed kernel time and the number of timed launches since dq_set_timing. */
int dq_set_timing(dq_state *s, int on);
int dq_get_timing(dq_state *s, double *sweep_ms, uint64_t *sweep_launches);

/* Amplitudes the state-sweep kernels read and wrote (each once per sweep)
 * since dq_set_timing (no reference counterpart; measurement only).  A sweep
 * visits only the tiles where qubits that are still 0 in every amplitude
 * (none touched since the last dq_state_reset) are 0, so this can be less
 * than sweeps x 2^n. */
int dq_get_timed_amplitudes(dq_state *s, uint64_t *amps);

/* Whether the engine uses the qubits it knows to be 0 in every amplitude
 * (default on; no reference counterpart): register sweeps skip the tiles
 * where such a qubit is 1, and dq_prob_one of such a qubit returns +0.0
 * without reading the state.  Those tiles hold zeros that a sweep would
 * leave zero, so either setting gives the same state; off (until the next
 * reset after switching on again) is for timing comparisons. */
int dq_set_zero_tiles(dq_state *s, int on);

/* Number of HBM sweeps and kernel launches issued so far on this handle. */
int dq_get_stats(const dq_state *s, uint64_t *sweeps, uint64_t *launches,
                 uint64_t *gates);

/* Pair fusion at fusion levels 2 and 3 (no reference counterpart; tuning
 * only; default on): neighbouring gates on at most 2 tile qubits are merged
 * into one op (their FP64 product, classified again). */
int dq_set_pair_fusion(dq_state *s, int on);

/* Register layout of the complex128 sweeps at fusion levels 2-3 (no
 * reference counterpart; tuning only): 4 (default) keeps 16 amplitudes per
 * thread in registers (k_tile_reg, 2 waves per SIMD), 3 keeps 8
 * (k_tile_reg8, 4 waves per SIMD, more layout changes).  Results agree to
 * rounding. */
int dq_set_reg_bits(dq_state *s, int reg_bits);

/* Fusion level 3 cost threshold (no reference counterpart; tuning only): a
 * block of gates is fused into one MFMA GEMM only when its gates' estimated
 * cost is at least min_cost FP32 flops per amplitude (default 128, about
 * the cost of the GEMM itself); 0 fuses every block of >= 3 qubits and >= 4
 * gates. */
int dq_set_fuse_min_cost(dq_state *s, double min_cost);

/* Fusion level 3: 5-qubit blocks applied on MFMA so far, and the gates they
 * replaced (no reference counterpart). */
int dq_get_fused_stats(const dq_state *s, uint64_t *blocks, uint64_t *gates);

/* ---- batching: many independent circuits in one call ----
 * No reference counterpart (the reference simulates one circuit per call;
 * its closest loop is the per-shot loop of
 * src/monolithic/statevector/simulator.rs:250-262).  Each instance is a
 * unitary circuit from |0...0> on its own n qubits plus a diagonal
 * observable H = sum_t coeff_t Z(zmask_t) (default: the open chain
 * sum_{i=0}^{n-2} Z_i Z_{i+1}); dq_batch_run writes <psi|H|psi> of every
 * instance.  Three paths:
 *   - LDS-resident: the state fits one workgroup's LDS (n <= 13 for
 *     complex128, n <= 14 for complex64): one workgroup per instance;
 *   - persistent: 14..20 qubits, one workgroup per instance in one launch
 *     that takes instances from a device queue, state in a global arena;
 *   - engine: larger instances on the statevector engine from `workers`
 *     host threads, each with its own states and stream.
 * The LDS and persistent work of a run is one captured hipGraph (upload,
 * counter reset, launches, download).  Above the LDS limit the path and the
 * engine's sweep configuration are chosen per (width, dtype) by an
 * autotuner that times a four-layer proxy of the first instance of each
 * width once per process and caches the choice. */
typedef struct dq_batch dq_batch;

int dq_batch_create(int dtype, dq_batch **out);
void dq_batch_destroy(dq_batch *b);
/* Appends an instance on n_qubits (1..34); *index receives its number. */
int dq_batch_add_instance(dq_batch *b, int n_qubits, int *index);
/* Appends a gate (2^k x 2^k row-major (re, im) matrix, qubits[0] = MSB, as
 * dq_apply_nq) to an instance. */
int dq_batch_apply(dq_batch *b, int index, const double *u, const int *qubits, int k);
int dq_batch_set_observable(dq_batch *b, int index, int n_terms, const uint64_t *zmask,
                            const double *coeff);
int dq_batch_set_workers(dq_batch *b, int workers);
/* Runs every instance; expect[i] = <H> of instance i.  Synchronous. */
int dq_batch_run(dq_batch *b, double *expect);
/* Instances on each path in the last run, and the LDS launches' GPU time. */
int dq_batch_stats(const dq_batch *b, uint64_t *lds_instances, uint64_t *engine_instances,
                   double *lds_ms);
/* Autotuning on (1, default) or off (0: widths up to 20 take the persistent
 * path, larger ones the engine with its defaults). */
int dq_batch_set_autotune(dq_batch *b, int on);
/* Last run: counts[3] instances on the LDS / persistent / engine paths,
 * *graph_ms the GPU time of the captured graph (LDS + persistent launches),
 * *engine_s the wall time of the engine path, bytes[3] the algorithmic HBM
 * bytes per path (one read and one write of the state per sweep, plus
 * the initial write and the final read; LDS: 0). */
int dq_batch_path_stats(const dq_batch *b, uint64_t *counts, double *graph_ms, double *engine_s, double *bytes);
/* The autotuner's cached choice for (n_qubits, dtype): path 1 persistent or
 * 2 engine, and the engine's fusion level, tile qubits and register bits.
 * DQ_ERR_ARG if that width has not been tuned in this process. */
int dq_batch_tuned(int n_qubits, int dtype, int *path, int *fusion, int *tile_bits, int *reg_bits);

/* ---- batched shots ----
 * Replaces the per-shot loop of StatevectorSimulator::simulate_shots
 * (src/monolithic/statevector/simulator.rs:215-322): a program of gates,
 * measurements, resets and conditionals (the fused instruction list of
 * src/types.rs:254-293, built by the caller) runs `shots` times from
 * |0...0>, shot i drawing from ChaCha8Rng::seed_from_u64(base_seed + i)
 * (wrapping).  bits_out receives, per shot, dq_shots_words() 64-bit words of
 * classical bits (bit c of the shot's words = cbit c, 0 if never written);
 * shot_seconds (optional) each shot's device time.  Gates apply as
 * apply_one_qubit_seq / apply_n_qubit_seq (src/engine.rs:265-333) and
 * measurements as measure_qubit_seq (:335-360), sequential p1 sum included,
 * on the one-workgroup-per-shot path (n <= 13). */
typedef struct dq_shots dq_shots;

int dq_shots_create(int n_qubits, int num_cbits, dq_shots **out);
void dq_shots_destroy(dq_shots *s);
int dq_shots_add_gate(dq_shots *s, const double *u, const int *qubits, int k);
int dq_shots_add_measure(dq_shots *s, int qubit, int cbit);
int dq_shots_add_reset(dq_shots *s, int qubit);
/* The next n_ops program ops run only if the classical bits
 * creg_base..creg_base+creg_size-1 (bit b weighted 2^b) equal creg_value
 * (simulator.rs:841-850). */
int dq_shots_add_conditional(dq_shots *s, int creg_base, int creg_size, uint64_t creg_value, int n_ops);
int dq_shots_words(const dq_shots *s);
int dq_shots_run(dq_shots *s, uint64_t base_seed, uint64_t shots, uint64_t *bits_out,
                 double *shot_seconds);

/* ---- matrix-product-state contraction ----
 * Replaces Mps::to_statevector and the p1 sum of Mps::measure
 * (src/monolithic/mps/simulator.rs:246-279); the site tensors themselves
 * stay with the caller (the reference's Mps of :68-74).  bond[0..n] are the
 * bond dimensions (bond[0] = bond[n] = 1, each <= 65536); site q is a
 * (bond[q], 2, bond[q+1]) array of complex128, row-major as Tensor::idx
 * (:49-52), interleaved (re, im); the sites are concatenated in order.
 * psi(x) = A_0[bit 0 of x] ... A_{n-1}[bit n-1 of x] (qubit q = index bit q). */
typedef struct dq_mps dq_mps;

int dq_mps_create(int n_sites, dq_mps **out);
void dq_mps_destroy(dq_mps *m);
/* Uploads the site tensors (a copy; later changes need another load). */
int dq_mps_load(dq_mps *m, const uint64_t *bond, const double *tensors);
/* Dense statevector of the loaded tensors into dst (dst has n_sites qubits;
 * any queued gates on dst are discarded).  Mps::to_statevector (:246-264). */
int dq_mps_to_state(dq_mps *m, dq_state *dst);
/* p1 = sum over x with bit `qubit` set of |psi(x)|^2, without writing the
 * statevector (the p1 of Mps::measure, :266-272, before its clamp). */
int dq_mps_prob_one(dq_mps *m, int qubit, double *p1);
/* Bond at which the chain is cut into the two contracted halves: -1 picks
 * it from the bond dimensions (default), 0..n forces it (for tests). */
int dq_mps_set_cut(dq_mps *m, int cut);
/* The cut the next contraction uses (-1 if nothing is loaded). */
int dq_mps_get_cut(dq_mps *m);

const char *dq_last_error(void);

#ifdef __cplusplus
}
#endif
#endif /* DQSIM_GPU_H */
