// mps.hpp -- host side of the matrix-product-state simulator.
//
// Restates the tensor network of src/monolithic/mps/simulator.rs:36-300:
// one (left, 2, right) tensor per qubit (Tensor, :36-66), single-qubit gates
// contracted into their site (apply_1q, :94-110), two-qubit gates brought
// next to each other by a SWAP network and applied to the merged pair
// followed by an SVD with truncation (apply_2q / apply_adjacent_2q,
// :112-244), and the projection of a measured site (project_qubit,
// :281-298).  The site tensors stay on the host, as in the reference; the
// dense contractions over all 2^n basis states (to_statevector and the p1 of
// measure) are the GPU's (dq_mps_* in include/dqsim_gpu.h).
//
// The SVD is a one-sided (Hestenes) Jacobi SVD written here from the
// textbook method (Demmel & Veselic, "Jacobi's method is more accurate than
// QR", SIAM J. Matrix Anal. Appl. 13 (1992)), with complex rotations; the
// reference calls nalgebra 0.33's DMatrix::svd, whose singular values come
// sorted in decreasing order.  Singular values agree to rounding; singular
// vectors agree up to the phase of each pair (and, for equal singular
// values, up to a rotation inside their subspace), which leaves the
// contracted state unchanged unless a truncation cuts through a group of
// equal singular values.
#pragma once
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <numeric>
#include <optional>
#include <stdexcept>
#include <vector>

#include "gates.hpp"

namespace dqh {

inline Cx cadd(Cx a, Cx b) { return {a.re + b.re, a.im + b.im}; }
inline Cx cconj(Cx a) { return {a.re, -a.im}; }
inline double cnorm2(Cx a) { return a.re * a.re + a.im * a.im; }

// simulator.rs:36-66: data[(left * 2 + state) * right + right_index]
struct MpsTensor {
    uint64_t left = 1, right = 1;
    std::vector<Cx> data;
    MpsTensor() = default;
    MpsTensor(uint64_t l, uint64_t r) : left(l), right(r), data(size_t(l * 2 * r), Cx{0.0, 0.0}) {}
    size_t idx(uint64_t l, uint64_t s, uint64_t r) const { return size_t((l * 2 + s) * right + r); }
};

// Thin SVD of the row-major m x n matrix a: a = U diag(s) Vt with U m x k,
// Vt k x n, k = min(m, n), s decreasing.  One-sided Jacobi on the columns of
// a (or of a^H when m < n): rotate column pairs until every pair is
// orthogonal to working precision; then s_j = |column j|, U_j = column j / s_j
// and V accumulates the rotations.
inline void svd_jacobi(int m, int n, const std::vector<Cx> &a, std::vector<Cx> &U, std::vector<double> &s,
                       std::vector<Cx> &Vt) {
    const bool tr = m < n;
    const int rows = tr ? n : m, cols = tr ? m : n;  // work matrix B: rows x cols, rows >= cols
    // column-major B so each column is contiguous
    std::vector<Cx> B(size_t(rows) * cols), V(size_t(cols) * cols, Cx{0.0, 0.0});
    for (int i = 0; i < m; ++i)
        for (int j = 0; j < n; ++j) {
            const Cx v = a[size_t(i) * n + j];
            if (!tr)
                B[size_t(j) * rows + i] = v;
            else
                B[size_t(i) * rows + j] = cconj(v);  // B = a^H: B[j][i] = conj(a[i][j])
        }
    for (int j = 0; j < cols; ++j) V[size_t(j) * cols + j] = Cx{1.0, 0.0};
    const double eps = std::max(rows, 4) * 2.220446049250313e-16;  // orthogonality reachable in FP64
    for (int sweep = 0; sweep < 80; ++sweep) {
        bool rotated = false;
        for (int p = 0; p < cols - 1; ++p)
            for (int q = p + 1; q < cols; ++q) {
                Cx *bp = &B[size_t(p) * rows], *bq = &B[size_t(q) * rows];
                double al = 0.0, be = 0.0;
                Cx g{0.0, 0.0};  // g = b_p^H b_q
                for (int i = 0; i < rows; ++i) {
                    al += cnorm2(bp[i]);
                    be += cnorm2(bq[i]);
                    g = cadd(g, cmul(cconj(bp[i]), bq[i]));
                }
                const double ga = std::hypot(g.re, g.im);
                if (ga == 0.0 || ga <= eps * std::sqrt(al * be)) continue;
                rotated = true;
                const double zeta = (be - al) / (2.0 * ga);
                const double t = (zeta >= 0.0 ? 1.0 : -1.0) / (std::fabs(zeta) + std::sqrt(1.0 + zeta * zeta));
                const double c = 1.0 / std::sqrt(1.0 + t * t), sn = c * t;
                const Cx e{g.re / ga, -g.im / ga};  // e^{-i phi}, g = |g| e^{i phi}
                // [b_p, b_q] <- [b_p, e b_q] [[c, s], [-s, c]]
                auto rot = [&](Cx *x, Cx *y, int len) {
                    for (int i = 0; i < len; ++i) {
                        const Cx ey = cmul(e, y[i]);
                        const Cx xp{c * x[i].re - sn * ey.re, c * x[i].im - sn * ey.im};
                        const Cx yp{sn * x[i].re + c * ey.re, sn * x[i].im + c * ey.im};
                        x[i] = xp;
                        y[i] = yp;
                    }
                };
                rot(bp, bq, rows);
                rot(&V[size_t(p) * cols], &V[size_t(q) * cols], cols);
            }
        if (!rotated) break;
    }
    std::vector<double> nrm(cols);
    for (int j = 0; j < cols; ++j) {
        double acc = 0.0;
        for (int i = 0; i < rows; ++i) acc += cnorm2(B[size_t(j) * rows + i]);
        nrm[j] = std::sqrt(acc);
    }
    std::vector<int> ord(cols);
    std::iota(ord.begin(), ord.end(), 0);
    std::stable_sort(ord.begin(), ord.end(), [&](int x, int y) { return nrm[x] > nrm[y]; });
    const int k = cols;
    s.assign(k, 0.0);
    // W_j = B_j / s_j (rows), V_j (cols): a = W S V^H when !tr; a^H = W S V^H -> a = V S W^H when tr.
    U.assign(size_t(m) * k, Cx{0.0, 0.0});
    Vt.assign(size_t(k) * n, Cx{0.0, 0.0});
    for (int jj = 0; jj < k; ++jj) {
        const int j = ord[jj];
        s[jj] = nrm[j];
        const double inv = nrm[j] > 0.0 ? 1.0 / nrm[j] : 0.0;
        for (int i = 0; i < rows; ++i) {
            const Cx w{B[size_t(j) * rows + i].re * inv, B[size_t(j) * rows + i].im * inv};
            if (!tr)
                U[size_t(i) * k + jj] = w;       // U = W (m x k)
            else
                Vt[size_t(jj) * n + i] = cconj(w);  // Vt = W^H (k x n)
        }
        for (int i = 0; i < cols; ++i) {
            const Cx v = V[size_t(j) * cols + i];
            if (!tr)
                Vt[size_t(jj) * n + i] = cconj(v);  // Vt = V^H
            else
                U[size_t(i) * k + jj] = v;         // U = V (m x k)
        }
    }
}

// simulator.rs:68-92 and the gate methods.
class MpsNetwork {
   public:
    MpsNetwork(int n, std::optional<uint64_t> max_bond, double threshold)
        : tensors(size_t(n)), max_bond_(max_bond), thr_(threshold) {
        for (auto &t : tensors) {
            t = MpsTensor(1, 1);
            t.data[t.idx(0, 0, 0)] = Cx{1.0, 0.0};
        }
    }

    std::vector<MpsTensor> tensors;
    uint64_t svd_calls = 0;
    double svd_time = 0.0;

    int num_sites() const { return int(tensors.size()); }

    // apply_1q, simulator.rs:94-110 (m: row-major 2x2)
    void apply_1q(int q, const Mat &m) {
        MpsTensor &t = tensors[size_t(q)];
        MpsTensor nt(t.left, t.right);
        for (uint64_t l = 0; l < t.left; ++l)
            for (uint64_t r = 0; r < t.right; ++r)
                for (int out = 0; out < 2; ++out) {
                    Cx acc{0.0, 0.0};
                    for (int in = 0; in < 2; ++in) acc = cadd(acc, cmul(m.get(out, in), t.data[t.idx(l, in, r)]));
                    nt.data[nt.idx(l, out, r)] = acc;
                }
        t = std::move(nt);
    }

    // apply_2q, simulator.rs:112-135: SWAP b next to a, apply, SWAP back.
    void apply_2q(int a, int b, const Mat &m) {
        if (a == b) return;
        const int lo = std::min(a, b), hi = std::max(a, b);
        Mat sw;
        gate_matrix("swap", {}, sw);
        for (int pos = hi - 1; pos > lo; --pos) apply_adjacent_2q(pos, sw);
        apply_adjacent_2q(lo, a < b ? m : reverse_2q(m));
        for (int pos = lo + 1; pos < hi; ++pos) apply_adjacent_2q(pos, sw);
    }

    // reverse_2q_order, simulator.rs:565-579: exchange the roles of the two qubits.
    static Mat reverse_2q(const Mat &m) {
        Mat out = identity(2);
        for (int r = 0; r < 4; ++r)
            for (int c = 0; c < 4; ++c) out.set(r, c, m.get(((r & 1) << 1) | (r >> 1), ((c & 1) << 1) | (c >> 1)));
        return out;
    }

    // apply_adjacent_2q, simulator.rs:137-244
    void apply_adjacent_2q(int q, const Mat &m) {
        const MpsTensor &A = tensors[size_t(q)], &Bt = tensors[size_t(q) + 1];
        if (A.right != Bt.left) throw std::runtime_error("Invalid MPS bond dimensions");
        const uint64_t L = A.left, K = A.right, R = Bt.right;
        // P[(l, i0), (i1, r)] = sum_k A[l][i0][k] B[k][i1][r]
        std::vector<Cx> P(size_t(L * 2 * 2 * R), Cx{0.0, 0.0});
        for (uint64_t l = 0; l < L; ++l)
            for (int i0 = 0; i0 < 2; ++i0)
                for (uint64_t k = 0; k < K; ++k) {
                    const Cx a = A.data[A.idx(l, uint64_t(i0), k)];
                    Cx *row = &P[size_t((l * 2 + uint64_t(i0)) * 2 * R)];
                    for (int i1 = 0; i1 < 2; ++i1)
                        for (uint64_t r = 0; r < R; ++r)
                            row[size_t(uint64_t(i1) * R + r)] =
                                cadd(row[size_t(uint64_t(i1) * R + r)], cmul(a, Bt.data[Bt.idx(k, uint64_t(i1), r)]));
                }
        // theta[(l, o0), (o1, r)] = sum_{i0, i1} m[o0 o1][i0 i1] P[(l, i0), (i1, r)]
        const int rows = int(L * 2), cols = int(2 * R);
        std::vector<Cx> theta(size_t(rows) * cols, Cx{0.0, 0.0});
        for (uint64_t l = 0; l < L; ++l)
            for (int o0 = 0; o0 < 2; ++o0)
                for (int o1 = 0; o1 < 2; ++o1)
                    for (uint64_t r = 0; r < R; ++r) {
                        Cx acc{0.0, 0.0};
                        for (int i0 = 0; i0 < 2; ++i0)
                            for (int i1 = 0; i1 < 2; ++i1)
                                acc = cadd(acc, cmul(m.get(o0 * 2 + o1, i0 * 2 + i1),
                                                     P[size_t((l * 2 + uint64_t(i0)) * 2 * R + uint64_t(i1) * R + r)]));
                        theta[size_t((l * 2 + uint64_t(o0)) * uint64_t(cols) + uint64_t(o1) * R + r)] = acc;
                    }
        const auto t0 = std::chrono::steady_clock::now();
        std::vector<Cx> U, Vt;
        std::vector<double> s;
        svd_jacobi(rows, cols, theta, U, s, Vt);
        svd_time += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        svd_calls++;
        const size_t ks = s.size();
        // simulator.rs:185-218: truncation and the norm-preserving rescale
        size_t keep = 0;
        for (double v : s)
            if (v > thr_) ++keep;
        keep = std::max<size_t>(keep, 1);
        if (max_bond_) keep = std::min<size_t>(keep, size_t(std::max<uint64_t>(*max_bond_, 1)));
        double kept = 0.0, total = 0.0, discarded = 0.0;
        for (size_t j = 0; j < keep && j < ks; ++j) kept += s[j] * s[j];
        for (size_t j = 0; j < ks; ++j) total += s[j] * s[j];
        for (size_t j = keep; j < ks; ++j) discarded += s[j] * s[j];
        kept = std::max(std::sqrt(kept), 1e-15);
        total = std::sqrt(total);
        const double scale = discarded > 0.0 ? total / kept : 1.0;
        keep = std::min(keep, ks);
        MpsTensor nl(L, keep), nr(keep, R);
        for (uint64_t l = 0; l < L; ++l)
            for (int st = 0; st < 2; ++st)
                for (size_t b = 0; b < keep; ++b)
                    nl.data[nl.idx(l, uint64_t(st), b)] = U[size_t((l * 2 + uint64_t(st)) * ks + b)];
        for (size_t b = 0; b < keep; ++b) {
            const double sig = s[b] * scale;
            for (int st = 0; st < 2; ++st)
                for (uint64_t r = 0; r < R; ++r) {
                    const Cx v = Vt[b * size_t(cols) + size_t(uint64_t(st) * R + r)];
                    nr.data[nr.idx(b, uint64_t(st), r)] = Cx{sig * v.re, sig * v.im};
                }
        }
        tensors[size_t(q)] = std::move(nl);
        tensors[size_t(q) + 1] = std::move(nr);
    }

    // project_qubit, simulator.rs:281-298
    void project(int q, int outcome, double prob) {
        const double scale = prob > 0.0 ? 1.0 / std::sqrt(prob) : 0.0;
        MpsTensor &t = tensors[size_t(q)];
        for (uint64_t l = 0; l < t.left; ++l)
            for (int st = 0; st < 2; ++st)
                for (uint64_t r = 0; r < t.right; ++r) {
                    Cx &v = t.data[t.idx(l, uint64_t(st), r)];
                    v = st == outcome ? Cx{v.re * scale, v.im * scale} : Cx{0.0, 0.0};
                }
    }

    // Bond dimensions and the concatenated tensors in the layout dq_mps_load takes.
    void pack(std::vector<uint64_t> &bond, std::vector<double> &flat) const {
        const size_t n = tensors.size();
        bond.assign(n + 1, 1);
        size_t total = 0;
        for (size_t q = 0; q < n; ++q) {
            bond[q] = tensors[q].left;
            bond[q + 1] = tensors[q].right;
            total += tensors[q].data.size();
        }
        flat.resize(2 * total);
        size_t o = 0;
        for (const auto &t : tensors)
            for (const Cx &v : t.data) {
                flat[o++] = v.re;
                flat[o++] = v.im;
            }
    }

   private:
    std::optional<uint64_t> max_bond_;
    double thr_;
};

}  // namespace dqh
