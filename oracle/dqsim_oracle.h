/*
 * dqsim_oracle.h -- CPU restatement of the reference statevector engine.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (dqsim_bosonic_amd/) links,
 * loads or calls this code.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py use it, and only as the checker / the timed
 * CPU baseline.
 *
 * State layout: interleaved (re, im) doubles, amplitude index bit q <-> qubit q
 * (qubit 0 = LSB), as src/engine.rs:34,113.  Multi-qubit matrices are
 * row-major 2^k x 2^k with qubits[0] the MSB of the row/column index
 * (src/engine.rs:104-106,116-125).
 *
 * Every function names the reference file:line it restates.  Floating-point
 * operations follow num-complex 0.4.6 op order with no contraction
 * (compile with -ffp-contract=off):
 *   mul (ar*br - ai*bi, ar*bi + ai*br), add componentwise,
 *   /= f64 divides both parts, norm_sqr = re*re + im*im.
 */
#ifndef DQSIM_ORACLE_H
#define DQSIM_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- ChaCha8Rng (rand_chacha 0.3.1) + seed_from_u64 (rand_core 0.6.4) ---- */
typedef struct {
    uint32_t key[8];
    uint64_t block_counter; /* next block to generate */
    uint32_t buf[64];       /* 4 blocks, as BlockRng<ChaCha8Core> results */
    int idx;                /* next word index; 64 = empty */
    int rounds;             /* 8 for ChaCha8 */
} or_rng;

void or_chacha_block(const uint32_t key[8], uint64_t counter, uint64_t nonce,
                     int rounds, uint32_t out[16]);
uint32_t or_pcg32_output(uint64_t state);
void or_rng_from_seed(or_rng *r, const uint8_t seed[32], int rounds);
void or_rng_seed_from_u64(or_rng *r, uint64_t seed);
uint32_t or_rng_next_u32(or_rng *r);
uint64_t or_rng_next_u64(or_rng *r);
double or_rng_gen_f64(or_rng *r);

/* ---- engine.rs ---- */
void or_apply_one_qubit(double *state, int n, const double u[8], int target,
                        const int *ctrl_wires, const int *ctrl_vals, int n_ctrl);
void or_apply_n_qubit(double *state, int n, const double *u, const int *qubits,
                      int k);
double or_prob_one(const double *state, int n, int qubit);
void or_collapse(double *state, int n, int qubit, int outcome, double p1);
int or_measure_qubit(double *state, int n, int qubit, or_rng *rng, double *p1_out);
void or_marginal_probs(const double *state, int n, const int *qubits, int k,
                       double *out);
void or_marginal_probs_literal(const double *state, int n, const int *qubits,
                               int k, double *out);
void or_sample_counts(const double *probs, int k, uint64_t shots, or_rng *rng,
                      uint32_t *idx_out);
void or_inner(const double *a, const double *b, uint64_t len, double *re,
              double *im);

/* Mps::to_statevector (src/monolithic/mps/simulator.rs:246-264) for the basis
 * indices x_begin .. x_begin + count - 1 (out: count complex values). */
void or_mps_amplitudes(int n, const uint64_t *bond, const double *tensors, uint64_t x_begin, uint64_t count,
                       double *out);

/* OpenMP variants with the reference's Rayon decomposition (CPU baseline). */
void or_apply_one_qubit_par(double *state, int n, const double u[8], int target);
void or_apply_n_qubit_par(double *state, int n, const double *u,
                          const int *qubits, int k);
double or_prob_one_par(const double *state, int n, int qubit);
void or_collapse_par(double *state, int n, int qubit, int outcome, double p1);
int or_num_threads(void);

/* ---- gates.rs ---- */
/* Fills out[(2^k)^2 * 2] for gate `name` with params; returns k (arity) or -1. */
int or_gate_matrix(const char *name, const double *params, int n_params,
                   double *out);

#ifdef __cplusplus
}
#endif
#endif
