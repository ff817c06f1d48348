/*
 * dqsim_oracle.c -- CPU restatement of the reference statevector engine.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP engine and the CPU
 * baseline of the benchmark.  The product never links or calls it.
 *
 * Restates (reference paths relative to the reference repository):
 *   src/engine.rs:24-433       apply_one_qubit, apply_n_qubit, measure_qubit,
 *                              marginal_probs, sample_counts
 *   src/gates.rs:11-258        gate matrices
 *   src/monolithic/statevector/simulator.rs:184-191   fidelity inner product
 *   rand_core 0.6.4 SeedableRng::seed_from_u64 (PCG32 seed fill),
 *   rand_chacha 0.3.1 ChaCha8Rng (djb ChaCha, 64-bit counter, 4-block buffer),
 *   rand 0.8.5 Standard f64 = (next_u64 >> 11) * 2^-53.
 *   (Third-party crates pinned in the reference's Cargo.lock:376-400; the
 *   algorithms are restated from their published definitions.)
 *
 * Must be compiled with -ffp-contract=off and without -ffast-math so that
 * every complex product/sum rounds exactly like num-complex 0.4.6 on the
 * reference (LLVM's default for Rust is no contraction).
 */
#include "dqsim_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef struct { double re, im; } cx;

static inline cx C(double re, double im) { cx r; r.re = re; r.im = im; return r; }
/* num-complex Mul: (a.re*b.re - a.im*b.im, a.re*b.im + a.im*b.re) */
static inline cx cmul(cx a, cx b) {
    return C(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re);
}
static inline cx cadd(cx a, cx b) { return C(a.re + b.re, a.im + b.im); }
static inline cx cneg(cx a) { return C(-a.re, -a.im); }
static inline double nsq(cx a) { return a.re * a.re + a.im * a.im; }

static inline cx ld(const double *s, uint64_t i) { return C(s[2 * i], s[2 * i + 1]); }
static inline void st(double *s, uint64_t i, cx v) { s[2 * i] = v.re; s[2 * i + 1] = v.im; }

/* ====================================================================== */
/* ChaCha (djb layout: words 12-13 = 64-bit block counter, 14-15 = nonce)   */
/* ====================================================================== */
static inline uint32_t rotl32(uint32_t x, int r) { return (x << r) | (x >> (32 - r)); }
#define QR(a, b, c, d)                     \
    a += b; d ^= a; d = rotl32(d, 16);     \
    c += d; b ^= c; b = rotl32(b, 12);     \
    a += b; d ^= a; d = rotl32(d, 8);      \
    c += d; b ^= c; b = rotl32(b, 7);

void or_chacha_block(const uint32_t key[8], uint64_t counter, uint64_t nonce,
                     int rounds, uint32_t out[16]) {
    uint32_t in[16];
    in[0] = 0x61707865u; in[1] = 0x3320646eu; in[2] = 0x79622d32u; in[3] = 0x6b206574u;
    for (int i = 0; i < 8; ++i) in[4 + i] = key[i];
    in[12] = (uint32_t)counter; in[13] = (uint32_t)(counter >> 32);
    in[14] = (uint32_t)nonce;   in[15] = (uint32_t)(nonce >> 32);
    uint32_t x[16];
    memcpy(x, in, sizeof x);
    for (int r = 0; r < rounds; r += 2) {
        QR(x[0], x[4], x[8], x[12]);
        QR(x[1], x[5], x[9], x[13]);
        QR(x[2], x[6], x[10], x[14]);
        QR(x[3], x[7], x[11], x[15]);
        QR(x[0], x[5], x[10], x[15]);
        QR(x[1], x[6], x[11], x[12]);
        QR(x[2], x[7], x[8], x[13]);
        QR(x[3], x[4], x[9], x[14]);
    }
    for (int i = 0; i < 16; ++i) out[i] = x[i] + in[i];
}

void or_rng_from_seed(or_rng *r, const uint8_t seed[32], int rounds) {
    for (int i = 0; i < 8; ++i)
        r->key[i] = (uint32_t)seed[4 * i] | ((uint32_t)seed[4 * i + 1] << 8) |
                    ((uint32_t)seed[4 * i + 2] << 16) | ((uint32_t)seed[4 * i + 3] << 24);
    r->block_counter = 0;
    r->idx = 64;
    r->rounds = rounds;
}

/* rand_core 0.6 SeedableRng::seed_from_u64: PCG32 (XSH-RR) fills the seed in
 * 4-byte little-endian chunks, advancing the LCG before each output. */
/* PCG XSH-RR output permutation of a 64-bit LCG state (O'Neill 2014). */
uint32_t or_pcg32_output(uint64_t state) {
    uint32_t xs = (uint32_t)(((state >> 18) ^ state) >> 27);
    uint32_t rot = (uint32_t)(state >> 59);
    return (xs >> rot) | (xs << ((32 - rot) & 31));
}

void or_rng_seed_from_u64(or_rng *r, uint64_t state) {
    const uint64_t MUL = 6364136223846793005ull, INC = 11634580027462260723ull;
    uint8_t seed[32];
    for (int c = 0; c < 8; ++c) {
        state = state * MUL + INC;
        uint32_t x = or_pcg32_output(state);
        seed[4 * c] = (uint8_t)x; seed[4 * c + 1] = (uint8_t)(x >> 8);
        seed[4 * c + 2] = (uint8_t)(x >> 16); seed[4 * c + 3] = (uint8_t)(x >> 24);
    }
    or_rng_from_seed(r, seed, 8);
}

static void or_rng_refill(or_rng *r) {
    for (int b = 0; b < 4; ++b)
        or_chacha_block(r->key, r->block_counter + (uint64_t)b, 0, r->rounds, r->buf + 16 * b);
    r->block_counter += 4;
    r->idx = 0;
}

uint32_t or_rng_next_u32(or_rng *r) {
    if (r->idx >= 64) or_rng_refill(r);
    return r->buf[r->idx++];
}

/* rand_core BlockRng::next_u64: low word first; a draw straddling the
 * buffer end takes the last word, refills, then takes word 0. */
uint64_t or_rng_next_u64(or_rng *r) {
    if (r->idx < 63) {
        uint64_t lo = r->buf[r->idx], hi = r->buf[r->idx + 1];
        r->idx += 2;
        return lo | (hi << 32);
    }
    if (r->idx >= 64) {
        or_rng_refill(r);
        r->idx = 2;
        return (uint64_t)r->buf[0] | ((uint64_t)r->buf[1] << 32);
    }
    uint64_t lo = r->buf[63];
    or_rng_refill(r);
    r->idx = 1;
    return lo | ((uint64_t)r->buf[0] << 32);
}

double or_rng_gen_f64(or_rng *r) {
    return (double)(or_rng_next_u64(r) >> 11) * (1.0 / 9007199254740992.0);
}

/* ====================================================================== */
/* engine.rs                                                               */
/* ====================================================================== */

/* apply_one_qubit sequential branch, src/engine.rs:88-100 (and the
 * identical _seq twin :265-292).  Controls: (i & incl) == desired. */
void or_apply_one_qubit(double *state, int n, const double u[8], int target,
                        const int *ctrl_wires, const int *ctrl_vals, int n_ctrl) {
    uint64_t incl = 0, desired = 0;
    for (int c = 0; c < n_ctrl; ++c) {
        incl |= 1ull << ctrl_wires[c];
        if (ctrl_vals[c]) desired |= 1ull << ctrl_wires[c];
    }
    const uint64_t half = 1ull << target, dim = 1ull << n;
    const cx u00 = C(u[0], u[1]), u01 = C(u[2], u[3]);
    const cx u10 = C(u[4], u[5]), u11 = C(u[6], u[7]);
    for (uint64_t i = 0; i < dim; ++i) {
        if ((i & half) == 0 && (incl == 0 || (i & incl) == desired)) {
            uint64_t j = i | half;
            cx a = ld(state, i), b = ld(state, j);
            st(state, i, cadd(cmul(u00, a), cmul(u01, b)));
            st(state, j, cadd(cmul(u10, a), cmul(u11, b)));
        }
    }
}

/* offsets[j] for qubits[0] = MSB, src/engine.rs:116-125 */
static void nq_offsets(const int *qubits, int k, uint64_t *off) {
    int dim = 1 << k;
    for (int j = 0; j < dim; ++j) {
        uint64_t o = 0;
        for (int bit = 0; bit < k; ++bit)
            if ((j >> (k - 1 - bit)) & 1) o |= 1ull << qubits[bit];
        off[j] = o;
    }
}

/* dense update of one base, src/engine.rs:160-179: acc = 0; acc += U[r][c]*v[c] */
static inline void nq_base(double *state, const double *u, int dim,
                           const uint64_t *off, uint64_t base) {
    cx v[32], w[32];
    for (int j = 0; j < dim; ++j) v[j] = ld(state, base + off[j]);
    for (int r = 0; r < dim; ++r) {
        cx acc = C(0.0, 0.0);
        for (int c = 0; c < dim; ++c) {
            cx m = C(u[2 * (r * dim + c)], u[2 * (r * dim + c) + 1]);
            acc = cadd(acc, cmul(m, v[c]));
        }
        w[r] = acc;
    }
    for (int j = 0; j < dim; ++j) st(state, base + off[j], w[j]);
}

/* apply_n_qubit sequential branch, src/engine.rs:109-181 */
void or_apply_n_qubit(double *state, int n, const double *u, const int *qubits, int k) {
    const int dim = 1 << k;
    const uint64_t ns = 1ull << n;
    uint64_t mask = 0, off[32];
    for (int b = 0; b < k; ++b) mask |= 1ull << qubits[b];
    nq_offsets(qubits, k, off);
    for (uint64_t base = 0; base < ns; ++base)
        if ((base & mask) == 0) nq_base(state, u, dim, off, base);
}

/* p1 of measure_qubit, src/engine.rs:197-208 (sequential ascending sum). */
double or_prob_one(const double *state, int n, int qubit) {
    const uint64_t ns = 1ull << n, bit = 1ull << qubit;
    double p1 = 0.0;
    for (uint64_t i = 0; i < ns; ++i)
        if (i & bit) p1 += nsq(ld(state, i));
    return p1;
}

/* collapse of measure_qubit, src/engine.rs:212-254: zero the other half,
 * divide the kept half by max(sqrt(p1),1e-15) or max(sqrt(max(1-p1,0)),1e-15). */
void or_collapse(double *state, int n, int qubit, int outcome, double p1) {
    const uint64_t ns = 1ull << n, bit = 1ull << qubit;
    double norm = outcome ? fmax(sqrt(p1), 1e-15) : fmax(sqrt(fmax(1.0 - p1, 0.0)), 1e-15);
    for (uint64_t i = 0; i < ns; ++i) {
        int b = (i & bit) != 0;
        if (b != outcome) {
            st(state, i, C(0.0, 0.0));
        } else {
            state[2 * i] /= norm;
            state[2 * i + 1] /= norm;
        }
    }
}

/* measure_qubit, src/engine.rs:193-257: one f64 draw, outcome = r < p1. */
int or_measure_qubit(double *state, int n, int qubit, or_rng *rng, double *p1_out) {
    double p1 = or_prob_one(state, n, qubit);
    int outcome = or_rng_gen_f64(rng) < p1 ? 1 : 0;
    or_collapse(state, n, qubit, outcome, p1);
    if (p1_out) *p1_out = p1;
    return outcome;
}

/* marginal_probs literal restatement, src/engine.rs:368-390: for each j,
 * ascending filtered sum.  O(2^k * 2^n). */
void or_marginal_probs_literal(const double *state, int n, const int *qubits, int k,
                               double *out) {
    const uint64_t ns = 1ull << n, dim = 1ull << k;
    uint64_t mask = 0;
    for (int b = 0; b < k; ++b) mask |= 1ull << qubits[b];
    for (uint64_t j = 0; j < dim; ++j) {
        uint64_t desired = 0;
        for (int b = 0; b < k; ++b)
            if ((j >> (k - 1 - b)) & 1) desired |= 1ull << qubits[b];
        double p = 0.0;
        for (uint64_t i = 0; i < ns; ++i)
            if ((i & mask) == desired) p += nsq(ld(state, i));
        out[j] = p;
    }
}

/* Same values as or_marginal_probs_literal, bit for bit: one ascending pass
 * that adds |psi_i|^2 into its bin, so each bin sees its terms in the same
 * ascending order starting from 0.0.  O(2^n). */
void or_marginal_probs(const double *state, int n, const int *qubits, int k, double *out) {
    const uint64_t ns = 1ull << n, dim = 1ull << k;
    for (uint64_t j = 0; j < dim; ++j) out[j] = 0.0;
    for (uint64_t i = 0; i < ns; ++i) {
        uint64_t j = 0;
        for (int b = 0; b < k; ++b) j = (j << 1) | ((i >> qubits[b]) & 1);
        out[j] += nsq(ld(state, i));
    }
}

/* sample_counts, src/engine.rs:398-433: sequential CDF, last entry forced to
 * 1.0, idx = partition_point(c < r) clamped to 2^k-1, one f64 draw per shot.
 * Writes the sampled index of every shot (the caller histograms them). */
void or_sample_counts(const double *probs, int k, uint64_t shots, or_rng *rng,
                      uint32_t *idx_out) {
    const uint64_t dim = 1ull << k;
    double *cdf = (double *)malloc(sizeof(double) * dim);
    double acc = 0.0;
    for (uint64_t i = 0; i < dim; ++i) { acc += probs[i]; cdf[i] = acc; }
    cdf[dim - 1] = 1.0;
    for (uint64_t s = 0; s < shots; ++s) {
        double r = or_rng_gen_f64(rng);
        uint64_t lo = 0, hi = dim; /* first index with !(cdf < r) */
        while (lo < hi) {
            uint64_t mid = lo + (hi - lo) / 2;
            if (cdf[mid] < r) lo = mid + 1; else hi = mid;
        }
        if (lo > dim - 1) lo = dim - 1;
        idx_out[s] = (uint32_t)lo;
    }
    free(cdf);
}

/* fidelity's inner product, simulator.rs:184-191: sum of conj(a_i)*b_i folded
 * from 0 in index order. */
void or_inner(const double *a, const double *b, uint64_t len, double *re, double *im) {
    cx acc = C(0.0, 0.0);
    for (uint64_t i = 0; i < len; ++i) acc = cadd(acc, cmul(C(a[2 * i], -a[2 * i + 1]), ld(b, i)));
    *re = acc.re;
    *im = acc.im;
}

/* ---- OpenMP variants: the reference's Rayon decomposition (n >= 12) ---- */

int or_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* src/engine.rs:43-87: outer parallel loop over blocks of 2^(t+1); when a half
 * block holds >= 1024 amplitudes the pairs inside a block are split instead. */
void or_apply_one_qubit_par(double *state, int n, const double u[8], int target) {
    const uint64_t half = 1ull << target, block = half << 1, dim = 1ull << n;
    const uint64_t nblocks = dim / block;
    const cx u00 = C(u[0], u[1]), u01 = C(u[2], u[3]);
    const cx u10 = C(u[4], u[5]), u11 = C(u[6], u[7]);
    if (half >= 1024) {
        for (uint64_t b = 0; b < nblocks; ++b) {
            uint64_t base = b * block;
#pragma omp parallel for schedule(static)
            for (uint64_t k = 0; k < half; ++k) {
                uint64_t i = base + k, j = i + half;
                cx a = ld(state, i), bb = ld(state, j);
                st(state, i, cadd(cmul(u00, a), cmul(u01, bb)));
                st(state, j, cadd(cmul(u10, a), cmul(u11, bb)));
            }
        }
    } else {
#pragma omp parallel for schedule(static)
        for (uint64_t b = 0; b < nblocks; ++b) {
            uint64_t base = b * block;
            for (uint64_t k = 0; k < half; ++k) {
                uint64_t i = base + k, j = i + half;
                cx a = ld(state, i), bb = ld(state, j);
                st(state, i, cadd(cmul(u00, a), cmul(u01, bb)));
                st(state, j, cadd(cmul(u10, a), cmul(u11, bb)));
            }
        }
    }
}

/* src/engine.rs:127-156: parallel filter over all 2^n indices. */
void or_apply_n_qubit_par(double *state, int n, const double *u, const int *qubits, int k) {
    const int dim = 1 << k;
    const uint64_t ns = 1ull << n;
    uint64_t mask = 0, off[32];
    for (int b = 0; b < k; ++b) mask |= 1ull << qubits[b];
    nq_offsets(qubits, k, off);
#pragma omp parallel for schedule(static)
    for (uint64_t base = 0; base < ns; ++base)
        if ((base & mask) == 0) nq_base(state, u, dim, off, base);
}

/* src/engine.rs:197-202: per-block sums of the upper halves, then a total. */
double or_prob_one_par(const double *state, int n, int qubit) {
    const uint64_t bit = 1ull << qubit, block = bit << 1, nblocks = (1ull << n) / block;
    double p1 = 0.0;
#pragma omp parallel for reduction(+ : p1) schedule(static)
    for (uint64_t b = 0; b < nblocks; ++b) {
        double s = 0.0;
        for (uint64_t k = 0; k < bit; ++k) s += nsq(ld(state, b * block + bit + k));
        p1 += s;
    }
    return p1;
}

void or_collapse_par(double *state, int n, int qubit, int outcome, double p1) {
    const uint64_t ns = 1ull << n, bit = 1ull << qubit;
    double norm = outcome ? fmax(sqrt(p1), 1e-15) : fmax(sqrt(fmax(1.0 - p1, 0.0)), 1e-15);
#pragma omp parallel for schedule(static)
    for (uint64_t i = 0; i < ns; ++i) {
        int b = (i & bit) != 0;
        if (b != outcome) {
            st(state, i, C(0.0, 0.0));
        } else {
            state[2 * i] /= norm;
            state[2 * i + 1] /= norm;
        }
    }
}

/* ====================================================================== */
/* gates.rs                                                                */
/* ====================================================================== */
#define S2 0.70710678118654752440 /* FRAC_1_SQRT_2 */

static void put(double *out, int dim, int r, int c, cx v) {
    out[2 * (r * dim + c)] = v.re;
    out[2 * (r * dim + c) + 1] = v.im;
}
static void eye(double *out, int dim) {
    memset(out, 0, sizeof(double) * 2 * dim * dim);
    for (int i = 0; i < dim; ++i) out[2 * (i * dim + i)] = 1.0;
}
static void m2(double *out, cx a, cx b, cx c, cx d) {
    put(out, 2, 0, 0, a); put(out, 2, 0, 1, b); put(out, 2, 1, 0, c); put(out, 2, 1, 1, d);
}
/* u3, src/gates.rs:59-65 */
static void u3m(double th, double ph, double la, double *out) {
    double ct = cos(th / 2.0), s = sin(th / 2.0);
    m2(out, C(ct, 0.0), cneg(cmul(C(cos(la), sin(la)), C(s, 0.0))),
       cmul(C(cos(ph), sin(ph)), C(s, 0.0)), cmul(C(cos(ph + la), sin(ph + la)), C(ct, 0.0)));
}
/* controlled(U): rows/cols 2,3 hold U, src/gates.rs:121-129 */
static void ctrl4(const double *u2, double *out) {
    eye(out, 4);
    for (int r = 0; r < 2; ++r)
        for (int c = 0; c < 2; ++c) put(out, 4, 2 + r, 2 + c, C(u2[2 * (2 * r + c)], u2[2 * (2 * r + c) + 1]));
}

int or_gate_matrix(const char *name, const double *p, int np, double *out) {
    double t[8];
#define IS(s) (strcmp(name, s) == 0)
#define NEED(k) do { if (np < (k)) return -1; } while (0)
    /* 1q fixed, src/gates.rs:18-53 */
    if (IS("id") || IS("u0")) { eye(out, 2); return 1; }
    if (IS("x")) { m2(out, C(0, 0), C(1, 0), C(1, 0), C(0, 0)); return 1; }
    if (IS("y")) { m2(out, C(0, 0), C(0, -1), C(0, 1), C(0, 0)); return 1; }
    if (IS("z")) { m2(out, C(1, 0), C(0, 0), C(0, 0), C(-1, 0)); return 1; }
    if (IS("h")) { m2(out, C(S2, 0), C(S2, 0), C(S2, 0), C(-S2, 0)); return 1; }
    if (IS("s")) { m2(out, C(1, 0), C(0, 0), C(0, 0), C(0, 1)); return 1; }
    if (IS("sdg")) { m2(out, C(1, 0), C(0, 0), C(0, 0), C(0, -1)); return 1; }
    if (IS("t")) { m2(out, C(1, 0), C(0, 0), C(0, 0), C(S2, S2)); return 1; }
    if (IS("tdg")) { m2(out, C(1, 0), C(0, 0), C(0, 0), C(S2, -S2)); return 1; }
    if (IS("sx")) { m2(out, C(.5, .5), C(.5, -.5), C(.5, -.5), C(.5, .5)); return 1; }
    if (IS("sxdg")) { m2(out, C(.5, -.5), C(.5, .5), C(.5, .5), C(.5, -.5)); return 1; }
    /* 1q parametric, src/gates.rs:59-84 */
    if (IS("u3") || IS("u")) { NEED(3); u3m(p[0], p[1], p[2], out); return 1; }
    if (IS("u2")) { NEED(2); u3m(M_PI / 2.0, p[0], p[1], out); return 1; }
    if (IS("u1") || IS("p")) { NEED(1); m2(out, C(1, 0), C(0, 0), C(0, 0), C(cos(p[0]), sin(p[0]))); return 1; }
    if (IS("rx")) {
        NEED(1); double c = cos(p[0] / 2.0), s = sin(p[0] / 2.0);
        m2(out, C(c, 0), C(0, -s), C(0, -s), C(c, 0)); return 1;
    }
    if (IS("ry")) {
        NEED(1); double c = cos(p[0] / 2.0), s = sin(p[0] / 2.0);
        m2(out, C(c, 0), C(-s, 0), C(s, 0), C(c, 0)); return 1;
    }
    if (IS("rz")) {
        NEED(1);
        m2(out, C(cos(-p[0] / 2.0), sin(-p[0] / 2.0)), C(0, 0), C(0, 0), C(cos(p[0] / 2.0), sin(p[0] / 2.0)));
        return 1;
    }
    /* 2q fixed, src/gates.rs:90-115 */
    if (IS("cx") || IS("remote_cx")) { eye(out, 4); put(out, 4, 2, 2, C(0, 0)); put(out, 4, 3, 3, C(0, 0)); put(out, 4, 2, 3, C(1, 0)); put(out, 4, 3, 2, C(1, 0)); return 2; }
    if (IS("cz") || IS("nonlocal_cz") || IS("remote_cz")) { eye(out, 4); put(out, 4, 3, 3, C(-1, 0)); return 2; }
    if (IS("cy")) { eye(out, 4); put(out, 4, 2, 2, C(0, 0)); put(out, 4, 3, 3, C(0, 0)); put(out, 4, 2, 3, C(0, -1)); put(out, 4, 3, 2, C(0, 1)); return 2; }
    if (IS("ch")) { eye(out, 4); put(out, 4, 2, 2, C(S2, 0)); put(out, 4, 2, 3, C(S2, 0)); put(out, 4, 3, 2, C(S2, 0)); put(out, 4, 3, 3, C(-S2, 0)); return 2; }
    if (IS("swap")) { eye(out, 4); put(out, 4, 1, 1, C(0, 0)); put(out, 4, 2, 2, C(0, 0)); put(out, 4, 1, 2, C(1, 0)); put(out, 4, 2, 1, C(1, 0)); return 2; }
    if (IS("csx")) { eye(out, 4); put(out, 4, 2, 2, C(.5, .5)); put(out, 4, 2, 3, C(.5, -.5)); put(out, 4, 3, 2, C(.5, -.5)); put(out, 4, 3, 3, C(.5, .5)); return 2; }
    /* 2q parametric, src/gates.rs:131-160 */
    if (IS("crx")) { NEED(1); or_gate_matrix("rx", p, 1, t); ctrl4(t, out); return 2; }
    if (IS("cry")) { NEED(1); or_gate_matrix("ry", p, 1, t); ctrl4(t, out); return 2; }
    if (IS("crz")) { NEED(1); or_gate_matrix("rz", p, 1, t); ctrl4(t, out); return 2; }
    if (IS("cu1") || IS("cp")) { NEED(1); or_gate_matrix("u1", p, 1, t); ctrl4(t, out); return 2; }
    if (IS("cu3")) { NEED(3); u3m(p[0], p[1], p[2], t); ctrl4(t, out); return 2; }
    if (IS("cu")) {
        NEED(4); u3m(p[0], p[1], p[2], t);
        cx ph = C(cos(p[3]), sin(p[3]));
        for (int i = 0; i < 4; ++i) {
            cx v = cmul(ph, C(t[2 * i], t[2 * i + 1]));
            t[2 * i] = v.re; t[2 * i + 1] = v.im;
        }
        ctrl4(t, out); return 2;
    }
    if (IS("rxx")) {
        NEED(1); double c = cos(p[0] / 2.0), s = sin(p[0] / 2.0);
        memset(out, 0, sizeof(double) * 32);
        for (int i = 0; i < 4; ++i) { put(out, 4, i, i, C(c, 0)); put(out, 4, i, 3 - i, C(0, -s)); }
        return 2;
    }
    if (IS("rzz")) {
        NEED(1); double c = cos(p[0] / 2.0), s = sin(p[0] / 2.0);
        memset(out, 0, sizeof(double) * 32);
        put(out, 4, 0, 0, C(c, -s)); put(out, 4, 1, 1, C(c, s)); put(out, 4, 2, 2, C(c, s)); put(out, 4, 3, 3, C(c, -s));
        return 2;
    }
    /* remote-link gates, src/gates.rs:224-258 */
    if (IS("remote_link_psi_minus") || IS("remote_link_psi_plus")) {
        static const double pm[4][4] = {{0, S2, 0, S2}, {S2, 0, S2, 0}, {-S2, 0, S2, 0}, {0, -S2, 0, S2}};
        int plus = IS("remote_link_psi_plus");
        for (int r = 0; r < 4; ++r)
            for (int c = 0; c < 4; ++c) {
                cx v = C(pm[r][c], 0.0);
                if (plus && r >= 2) v = cneg(v);
                put(out, 4, r, c, v);
            }
        return 2;
    }
    if (IS("remote_link_phi_plus") || IS("remote_epr")) {
        static const double pp[4][4] = {{S2, 0, S2, 0}, {0, S2, 0, S2}, {0, S2, 0, -S2}, {S2, 0, -S2, 0}};
        for (int r = 0; r < 4; ++r)
            for (int c = 0; c < 4; ++c) put(out, 4, r, c, C(pp[r][c], 0.0));
        return 2;
    }
    /* 3q, src/gates.rs:166-183 */
    if (IS("ccx")) { eye(out, 8); put(out, 8, 6, 6, C(0, 0)); put(out, 8, 7, 7, C(0, 0)); put(out, 8, 6, 7, C(1, 0)); put(out, 8, 7, 6, C(1, 0)); return 3; }
    if (IS("cswap")) { eye(out, 8); put(out, 8, 5, 5, C(0, 0)); put(out, 8, 6, 6, C(0, 0)); put(out, 8, 5, 6, C(1, 0)); put(out, 8, 6, 5, C(1, 0)); return 3; }
    if (IS("rccx")) { eye(out, 8); put(out, 8, 6, 6, C(0, 0)); put(out, 8, 7, 7, C(0, 0)); put(out, 8, 6, 7, C(0, 1)); put(out, 8, 7, 6, C(0, -1)); return 3; }
    /* 4q, src/gates.rs:189-207 */
    if (IS("rc3x")) { eye(out, 16); put(out, 16, 14, 14, C(0, 1)); put(out, 16, 15, 15, C(0, -1)); return 4; }
    if (IS("c3x")) { eye(out, 16); put(out, 16, 14, 14, C(0, 0)); put(out, 16, 15, 15, C(0, 0)); put(out, 16, 14, 15, C(1, 0)); put(out, 16, 15, 14, C(1, 0)); return 4; }
    if (IS("c3sqrtx")) { eye(out, 16); put(out, 16, 14, 14, C(.5, .5)); put(out, 16, 14, 15, C(.5, -.5)); put(out, 16, 15, 14, C(.5, -.5)); put(out, 16, 15, 15, C(.5, .5)); return 4; }
    /* 5q, src/gates.rs:213-218 */
    if (IS("c4x")) { eye(out, 32); put(out, 32, 30, 30, C(0, 0)); put(out, 32, 31, 31, C(0, 0)); put(out, 32, 30, 31, C(1, 0)); put(out, 32, 31, 30, C(1, 0)); return 5; }
#undef IS
#undef NEED
    return -1;
}

/* Mps::to_statevector, src/monolithic/mps/simulator.rs:246-264, for the basis
 * indices x_begin .. x_begin + count - 1: per index, work = [1]; for each site
 * q in order, next[r] = sum over l ascending of work[l] * A_q[l][bit q][r];
 * amplitude = work[0].  bond[0..n] are the bond dimensions; site q is a
 * (bond[q], 2, bond[q+1]) row-major complex array, sites concatenated.
 * Single-threaded, as the reference's loop. */
void or_mps_amplitudes(int n, const uint64_t *bond, const double *tensors, uint64_t x_begin, uint64_t count,
                       double *out) {
    uint64_t maxb = 1, off[65];
    off[0] = 0;
    for (int q = 0; q < n; ++q) {
        if (bond[q + 1] > maxb) maxb = bond[q + 1];
        off[q + 1] = off[q] + 2 * bond[q] * bond[q + 1];
    }
    cx *work = (cx *)malloc(sizeof(cx) * maxb), *next = (cx *)malloc(sizeof(cx) * maxb);
    for (uint64_t t = 0; t < count; ++t) {
        const uint64_t x = x_begin + t;
        uint64_t wl = 1;
        work[0] = C(1.0, 0.0);
        for (int q = 0; q < n; ++q) {
            const uint64_t bit = (x >> q) & 1, L = bond[q], R = bond[q + 1];
            const double *A = tensors + 2 * off[q];
            for (uint64_t r = 0; r < R; ++r) next[r] = C(0.0, 0.0);
            for (uint64_t l = 0; l < wl && l < L; ++l)
                for (uint64_t r = 0; r < R; ++r)
                    next[r] = cadd(next[r], cmul(work[l], ld(A, (l * 2 + bit) * R + r)));
            cx *tmp = work;
            work = next;
            next = tmp;
            wl = R;
        }
        st(out, t, work[0]);
    }
    free(work);
    free(next);
}
