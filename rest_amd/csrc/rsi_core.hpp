// rsi_core.hpp — Gaussian integral core of the B producer (McMurchie–Davidson).
//
// Cartesian primitives x^i y^j z^k exp(-a r_A^2); products expanded in Hermite
// Gaussians (E coefficients), Coulomb-type integrals through the Hermite
// integrals R_tuv of the Boys function.  Used for the overlap, kinetic and
// nuclear-attraction one-electron matrices, the auxiliary Coulomb metric
// (P|Q) and the three-centre integrals (μν|P) that REST takes from libcint
// (src/molecule_io/mod.rs:700-760, :1072-1103, :1647).
#pragma once
#include <cmath>
#include <cstddef>
#include <vector>

namespace rsi {

constexpr int kMaxL = 4;                      // up to g functions
constexpr int kMaxLTot = 3 * kMaxL;           // R up to la+lb+lc (3c) or la+lb (2c)
constexpr int kBoysM = kMaxLTot + 1;          // Boys orders needed (+1 for kinetic safety)

inline int ncart(int l) { return (l + 1) * (l + 2) / 2; }

// Cartesian exponents of component c of shell l, ordered lx = l..0, ly = l-lx..0
inline void cart_exps(int l, int c, int* lx, int* ly, int* lz) {
    int idx = 0;
    for (int x = l; x >= 0; --x)
        for (int y = l - x; y >= 0; --y) {
            if (idx == c) {
                *lx = x;
                *ly = y;
                *lz = l - x - y;
                return;
            }
            ++idx;
        }
}

// ---------------------------------------------------------------------------
// Boys function F_m(T) = ∫_0^1 t^{2m} exp(-T t²) dt, m = 0..M.
// Small/medium T: 6-term Taylor expansion around a 0.05-spaced grid whose
// values come from the convergent series; large T: asymptotic F_0 and upward
// recursion (F_0 exact to < 1e-16 for T >= 36).
// ---------------------------------------------------------------------------
class Boys {
public:
    static constexpr int kOrd = 6;
    static constexpr double kStep = 0.05;
    static constexpr double kTMax = 36.0;
    static constexpr int kMTab = kBoysM + kOrd + 1;

    Boys() {
        n_ = (int)(kTMax / kStep) + 2;
        tab_.resize((size_t)n_ * kMTab);
        for (int i = 0; i < n_; ++i) series(i * kStep, kMTab - 1, &tab_[(size_t)i * kMTab]);
    }
    // F_0..F_M at T (M <= kBoysM)
    void eval(double T, int M, double* F) const {
        if (T < kTMax) {
            const int i = (int)(T / kStep + 0.5);
            const double dt = i * kStep - T;   // expand F_m(T) = Σ_k F_{m+k}(T_i) (T_i - T)^k / k!
            const double* row = &tab_[(size_t)i * kMTab];
            for (int m = 0; m <= M; ++m) {
                double s = 0.0, p = 1.0;
                for (int k = 0; k <= kOrd; ++k) {
                    s += row[m + k] * p;
                    p *= dt / (k + 1);
                }
                F[m] = s;
            }
        } else {
            const double e = std::exp(-T);
            F[0] = 0.5 * std::sqrt(M_PI / T);
            for (int m = 0; m < M; ++m) F[m + 1] = ((2 * m + 1) * F[m] - e) / (2.0 * T);
        }
    }
    // reference evaluation (series + downward recursion), F_0..F_M
    static void series(double T, int M, double* F) {
        const double e = std::exp(-T);
        if (T < 1e-15) {
            for (int m = 0; m <= M; ++m) F[m] = 1.0 / (2 * m + 1);
            return;
        }
        double term = 1.0 / (2 * M + 1), sum = term;
        for (int k = 1; k < 2000; ++k) {
            term *= 2.0 * T / (2 * M + 2 * k + 1);
            sum += term;
            if (term < 1e-18 * sum) break;
        }
        F[M] = e * sum;
        for (int m = M - 1; m >= 0; --m) F[m] = (2.0 * T * F[m + 1] + e) / (2 * m + 1);
    }

private:
    int n_ = 0;
    std::vector<double> tab_;
};

const Boys& boys();   // process-wide table

// ---------------------------------------------------------------------------
// Hermite expansion of a 1-D product x_A^i x_B^j exp(-a x_A² - b x_B²):
//   E[i][j][t], t = 0..i+j, without the exp(-μ X_AB²) prefactor.
// ---------------------------------------------------------------------------
struct HermiteE1 {
    // E indexed [i][j][t] with i <= li, j <= lj, t <= i+j
    static constexpr int N = kMaxL + 3;  // kinetic needs j up to lb+2
    double e[N][N][2 * N];
    void build(int li, int lj, double p, double XPA, double XPB) {
        const double oo2p = 0.5 / p;
        for (int i = 0; i <= li; ++i)
            for (int j = 0; j <= lj; ++j)
                for (int t = 0; t <= i + j + 1; ++t) e[i][j][t] = 0.0;
        e[0][0][0] = 1.0;
        for (int i = 0; i <= li; ++i) {
            for (int j = 0; j <= lj; ++j) {
                if (i == 0 && j == 0) continue;
                if (j == 0) {
                    // raise i
                    for (int t = 0; t <= i + j; ++t) {
                        double v = XPA * get(i - 1, j, t);
                        if (t > 0) v += oo2p * get(i - 1, j, t - 1);
                        v += (t + 1) * get(i - 1, j, t + 1);
                        e[i][j][t] = v;
                    }
                } else {
                    for (int t = 0; t <= i + j; ++t) {
                        double v = XPB * get(i, j - 1, t);
                        if (t > 0) v += oo2p * get(i, j - 1, t - 1);
                        v += (t + 1) * get(i, j - 1, t + 1);
                        e[i][j][t] = v;
                    }
                }
            }
        }
    }
    double get(int i, int j, int t) const { return (t < 0 || t > i + j) ? 0.0 : e[i][j][t]; }
};

// Hermite expansion of one Gaussian x^k exp(-g x²) about its own centre:
//   Ec[k][τ], τ = 0..k (single-centre special case, XPA = 0, no prefactor).
struct HermiteE0 {
    double e[kMaxL + 1][kMaxL + 1];
    void build(int l, double g) {
        const double oo2g = 0.5 / g;
        for (int k = 0; k <= l; ++k)
            for (int t = 0; t <= kMaxL; ++t) e[k][t] = 0.0;
        e[0][0] = 1.0;
        for (int k = 1; k <= l; ++k)
            for (int t = 0; t <= k; ++t) {
                double v = 0.0;
                if (t > 0) v += oo2g * e[k - 1][t - 1];
                if (t + 1 <= k - 1) v += (t + 1) * e[k - 1][t + 1];
                e[k][t] = v;
            }
    }
};

// ---------------------------------------------------------------------------
// Hermite Coulomb integrals R_tuv(α, X, Y, Z) for t+u+v <= L (n = 0 level),
// flat index r[(t*(L+1) + u)*(L+1) + v].
// ---------------------------------------------------------------------------
struct HermiteR {
    static constexpr int N = kMaxLTot + 1;
    int L = 0;
    double r[N * N * N];
    void build(int Lmax, double alpha, double X, double Y, double Z) {
        L = Lmax;
        const int n1 = L + 1;
        double F[kBoysM + 1];
        boys().eval(alpha * (X * X + Y * Y + Z * Z), L, F);
        // scratch R^{(n)}_{tuv} for all n via stacked levels: level n holds t+u+v <= L-n
        static thread_local std::vector<double> work;
        const size_t lvl = (size_t)n1 * n1 * n1;
        if (work.size() < lvl * n1) work.resize(lvl * n1);   // every entry read below is written first
        auto at = [&](int n, int t, int u, int v) -> double& {
            return work[(size_t)n * lvl + ((size_t)t * n1 + u) * n1 + v];
        };
        double fac = 1.0;
        for (int n = 0; n <= L; ++n) {
            at(n, 0, 0, 0) = fac * F[n];
            fac *= -2.0 * alpha;
        }
        for (int n = L - 1; n >= 0; --n) {
            const int top = L - n;
            for (int s = 1; s <= top; ++s)            // s = t+u+v at this level
                for (int t = 0; t <= s; ++t)
                    for (int u = 0; u <= s - t; ++u) {
                        const int v = s - t - u;
                        double val;
                        if (t > 0) {
                            val = X * at(n + 1, t - 1, u, v);
                            if (t > 1) val += (t - 1) * at(n + 1, t - 2, u, v);
                        } else if (u > 0) {
                            val = Y * at(n + 1, t, u - 1, v);
                            if (u > 1) val += (u - 1) * at(n + 1, t, u - 2, v);
                        } else {
                            val = Z * at(n + 1, t, u, v - 1);
                            if (v > 1) val += (v - 1) * at(n + 1, t, u, v - 2);
                        }
                        at(n, t, u, v) = val;
                    }
        }
        for (int t = 0; t <= L; ++t)
            for (int u = 0; u <= L - t; ++u)
                for (int v = 0; v <= L - t - u; ++v) r[(t * n1 + u) * n1 + v] = at(0, t, u, v);
    }
    double get(int t, int u, int v) const { return r[(t * (L + 1) + u) * (L + 1) + v]; }
};

// ---------------------------------------------------------------------------
// Cartesian -> real solid harmonic transformation (unnormalised rows; the
// basis layer rescales every row to a unit-norm function).  Order of m:
// -l..l.  Coefficients from the expansion of S_lm in x, y, z:
//   S_lm ∝ Σ_{t,u,v} (-1)^{t+v-vm} (1/4)^t C(l,t) C(l-t,|m|+t) C(t,u) C(|m|,2v)
//          x^{2t+|m|-2(u+v)} y^{2(u+v)} z^{l-2t-|m|}
// with vm = 0 (m >= 0) or 1/2 (m < 0) and v stepping by 1 from vm.
// ---------------------------------------------------------------------------
std::vector<double> cart2sph(int l);   // (2l+1) × ncart(l), row-major

}  // namespace rsi
