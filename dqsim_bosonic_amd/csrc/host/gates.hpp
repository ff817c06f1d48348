// gates.hpp -- host gate-matrix library.
//
// Closed-form FP64 matrices of the reference's src/gates.rs:18-258, built with
// the same libm cos/sin calls and the same complex products (num-complex op
// order, no contraction), so every entry has the reference's bits.
// Row-major, (re, im) interleaved; for multi-qubit gates the first-listed
// qubit is the MSB of the row/column index (src/gates.rs:87).
#pragma once
#include <cmath>
#include <string>
#include <vector>

namespace dqh {

struct Cx {
    double re, im;
};
inline Cx cmul(Cx a, Cx b) { return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
inline Cx cneg(Cx a) { return {-a.re, -a.im}; }

struct Mat {
    int k = 0;                 // arity
    std::vector<double> v;     // 2 * (2^k)^2 doubles
    int dim() const { return 1 << k; }
    void set(int r, int c, Cx x) {
        v[2 * (size_t(r) * dim() + c)] = x.re;
        v[2 * (size_t(r) * dim() + c) + 1] = x.im;
    }
    Cx get(int r, int c) const { return {v[2 * (size_t(r) * dim() + c)], v[2 * (size_t(r) * dim() + c) + 1]}; }
};

inline Mat identity(int k) {
    Mat m;
    m.k = k;
    m.v.assign(2 * size_t(1 << k) * (1 << k), 0.0);
    for (int i = 0; i < (1 << k); ++i) m.set(i, i, {1.0, 0.0});
    return m;
}

inline Mat mat2(Cx a, Cx b, Cx c, Cx d) {
    Mat m = identity(1);
    m.set(0, 0, a);
    m.set(0, 1, b);
    m.set(1, 0, c);
    m.set(1, 1, d);
    return m;
}

constexpr double kFrac1Sqrt2 = 0.70710678118654752440;  // std::f64::consts::FRAC_1_SQRT_2
constexpr double kPi = 3.14159265358979323846;

// src/gates.rs:59-65
inline Mat u3(double theta, double phi, double lam) {
    const double ct = std::cos(theta / 2.0), st = std::sin(theta / 2.0);
    return mat2({ct, 0.0}, cneg(cmul({std::cos(lam), std::sin(lam)}, {st, 0.0})),
                cmul({std::cos(phi), std::sin(phi)}, {st, 0.0}),
                cmul({std::cos(phi + lam), std::sin(phi + lam)}, {ct, 0.0}));
}
inline Mat u1(double lam) { return mat2({1, 0}, {0, 0}, {0, 0}, {std::cos(lam), std::sin(lam)}); }
inline Mat rx(double t) {
    const double c = std::cos(t / 2.0), s = std::sin(t / 2.0);
    return mat2({c, 0}, {0, -s}, {0, -s}, {c, 0});
}
inline Mat ry(double t) {
    const double c = std::cos(t / 2.0), s = std::sin(t / 2.0);
    return mat2({c, 0}, {-s, 0}, {s, 0}, {c, 0});
}
inline Mat rz(double p) {
    return mat2({std::cos(-p / 2.0), std::sin(-p / 2.0)}, {0, 0}, {0, 0}, {std::cos(p / 2.0), std::sin(p / 2.0)});
}
// controlled(U): U in rows/cols 2,3 (control = MSB), src/gates.rs:121-129
inline Mat controlled(const Mat &u) {
    Mat m = identity(2);
    for (int r = 0; r < 2; ++r)
        for (int c = 0; c < 2; ++c) m.set(2 + r, 2 + c, u.get(r, c));
    return m;
}
inline Mat swap_entries(int k, int a, int b, Cx ab, Cx ba, Cx aa = {0, 0}, Cx bb = {0, 0}) {
    Mat m = identity(k);
    m.set(a, a, aa);
    m.set(b, b, bb);
    m.set(a, b, ab);
    m.set(b, a, ba);
    return m;
}

// Builds the matrix of instruction kind / generic gate `name` (already
// mapped to its canonical kind) with the kind's parameters in field order.
// Returns false for an unknown name.
inline bool gate_matrix(const std::string &name, const std::vector<double> &p, Mat &out) {
    const double s2 = kFrac1Sqrt2;
    const Cx O{0, 0}, I{1, 0};
    if (name == "id" || name == "u0") { out = identity(1); return true; }
    if (name == "x") { out = mat2(O, I, I, O); return true; }
    if (name == "y") { out = mat2(O, {0, -1}, {0, 1}, O); return true; }
    if (name == "z") { out = mat2(I, O, O, {-1, 0}); return true; }
    if (name == "h") { out = mat2({s2, 0}, {s2, 0}, {s2, 0}, {-s2, 0}); return true; }
    if (name == "s") { out = mat2(I, O, O, {0, 1}); return true; }
    if (name == "sdg") { out = mat2(I, O, O, {0, -1}); return true; }
    if (name == "t") { out = mat2(I, O, O, {s2, s2}); return true; }
    if (name == "tdg") { out = mat2(I, O, O, {s2, -s2}); return true; }
    if (name == "sx") { out = mat2({.5, .5}, {.5, -.5}, {.5, -.5}, {.5, .5}); return true; }
    if (name == "sxdg") { out = mat2({.5, -.5}, {.5, .5}, {.5, .5}, {.5, -.5}); return true; }
    if (name == "u3" || name == "u") { out = u3(p.at(0), p.at(1), p.at(2)); return true; }
    if (name == "u2") { out = u3(kPi / 2.0, p.at(0), p.at(1)); return true; }
    if (name == "u1" || name == "p") { out = u1(p.at(0)); return true; }
    if (name == "rx") { out = rx(p.at(0)); return true; }
    if (name == "ry") { out = ry(p.at(0)); return true; }
    if (name == "rz") { out = rz(p.at(0)); return true; }
    // two-qubit fixed, src/gates.rs:90-115
    if (name == "cx") { out = swap_entries(2, 2, 3, I, I); return true; }
    if (name == "cz") { out = identity(2); out.set(3, 3, {-1, 0}); return true; }
    if (name == "cy") { out = swap_entries(2, 2, 3, {0, -1}, {0, 1}); return true; }
    if (name == "ch") { out = swap_entries(2, 2, 3, {s2, 0}, {s2, 0}, {s2, 0}, {-s2, 0}); return true; }
    if (name == "swap") { out = swap_entries(2, 1, 2, I, I); return true; }
    if (name == "csx") { out = swap_entries(2, 2, 3, {.5, -.5}, {.5, -.5}, {.5, .5}, {.5, .5}); return true; }
    // two-qubit parametric, src/gates.rs:131-160
    if (name == "crx") { out = controlled(rx(p.at(0))); return true; }
    if (name == "cry") { out = controlled(ry(p.at(0))); return true; }
    if (name == "crz") { out = controlled(rz(p.at(0))); return true; }
    if (name == "cu1" || name == "cp") { out = controlled(u1(p.at(0))); return true; }
    if (name == "cu3") { out = controlled(u3(p.at(0), p.at(1), p.at(2))); return true; }
    if (name == "cu") {
        const Cx ph{std::cos(p.at(3)), std::sin(p.at(3))};
        Mat inner = u3(p.at(0), p.at(1), p.at(2));
        for (int r = 0; r < 2; ++r)
            for (int c = 0; c < 2; ++c) inner.set(r, c, cmul(ph, inner.get(r, c)));
        out = controlled(inner);
        return true;
    }
    if (name == "rxx") {
        const double c = std::cos(p.at(0) / 2.0), s = std::sin(p.at(0) / 2.0);
        out = identity(2);
        for (int i = 0; i < 4; ++i) {
            out.set(i, i, {c, 0});
            out.set(i, 3 - i, {0, -s});
        }
        return true;
    }
    if (name == "rzz") {
        const double c = std::cos(p.at(0) / 2.0), s = std::sin(p.at(0) / 2.0);
        out = identity(2);
        out.set(0, 0, {c, -s});
        out.set(1, 1, {c, s});
        out.set(2, 2, {c, s});
        out.set(3, 3, {c, -s});
        return true;
    }
    // remote-link gates, src/gates.rs:224-258
    if (name == "remote_link_psi_minus" || name == "remote_link_psi_plus") {
        const double pm[4][4] = {{0, s2, 0, s2}, {s2, 0, s2, 0}, {-s2, 0, s2, 0}, {0, -s2, 0, s2}};
        const bool plus = name == "remote_link_psi_plus";
        out = identity(2);
        for (int r = 0; r < 4; ++r)
            for (int c = 0; c < 4; ++c) {
                Cx v{pm[r][c], 0.0};
                if (plus && r >= 2) v = cneg(v);  // (Z (x) I) . psi_minus
                out.set(r, c, v);
            }
        return true;
    }
    if (name == "remote_link_phi_plus") {
        const double pp[4][4] = {{s2, 0, s2, 0}, {0, s2, 0, s2}, {0, s2, 0, -s2}, {s2, 0, -s2, 0}};
        out = identity(2);
        for (int r = 0; r < 4; ++r)
            for (int c = 0; c < 4; ++c) out.set(r, c, {pp[r][c], 0.0});
        return true;
    }
    // three-, four-, five-qubit, src/gates.rs:166-218
    if (name == "ccx") { out = swap_entries(3, 6, 7, I, I); return true; }
    if (name == "cswap") { out = swap_entries(3, 5, 6, I, I); return true; }
    if (name == "rccx") { out = swap_entries(3, 6, 7, {0, 1}, {0, -1}); return true; }
    if (name == "rc3x") { out = identity(4); out.set(14, 14, {0, 1}); out.set(15, 15, {0, -1}); return true; }
    if (name == "c3x") { out = swap_entries(4, 14, 15, I, I); return true; }
    if (name == "c3sqrtx") {
        out = swap_entries(4, 14, 15, {.5, -.5}, {.5, -.5}, {.5, .5}, {.5, .5});
        return true;
    }
    if (name == "c4x") { out = swap_entries(5, 30, 31, I, I); return true; }
    return false;
}

}  // namespace dqh
