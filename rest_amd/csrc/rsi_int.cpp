// rsi_int.cpp — basis construction and host integral drivers of the B
// producer: one-electron S/T/V, the auxiliary metric (P|Q) and the
// three-centre integrals (μν|P) on spherical shells (McMurchie–Davidson,
// OpenMP over shell pairs).
#include <omp.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/rest_ri.h"
#include "rsi_core.hpp"
#include "rsi_internal.hpp"

namespace rsi {

const Boys& boys() {
    static const Boys b;
    return b;
}

namespace {
thread_local std::string g_err;

double binom(int n, int k) {
    if (k < 0 || k > n) return 0.0;
    double r = 1.0;
    for (int i = 1; i <= k; ++i) r = r * (n - k + i) / i;
    return r;
}

int cart_index(int l, int px, int py) {
    int idx = 0;
    for (int x = l; x >= 0; --x)
        for (int y = l - x; y >= 0; --y) {
            if (x == px && y == py) return idx;
            ++idx;
        }
    return -1;
}

// ∫ x^n exp(-p x²) dx over the real line
double gauss_moment(int n, double p) {
    if (n & 1) return 0.0;
    double df = 1.0;  // (n-1)!!
    for (int k = n - 1; k > 1; k -= 2) df *= k;
    return df * std::sqrt(M_PI / p) / std::pow(2.0 * p, n / 2);
}

// radial normalisation of r^l exp(-a r²):  1/sqrt(∫ r^{2l+2} exp(-2a r²) dr)
double gto_norm(int l, double a) {
    const double g = std::tgamma(l + 1.5) / (2.0 * std::pow(2.0 * a, l + 1.5));
    return 1.0 / std::sqrt(g);
}

}  // namespace

void set_error(const std::string& m) { g_err = m; }
const std::string& last_error() { return g_err; }

std::vector<double> cart2sph(int l) {
    const int nc = ncart(l), ns = 2 * l + 1;
    std::vector<double> T((size_t)ns * nc, 0.0);
    if (l == 0) {
        T[0] = 1.0;
        return T;
    }
    if (l == 1) {  // libcint order for p shells: x, y, z (= the Cartesian order)
        for (int i = 0; i < 3; ++i) T[i * 3 + i] = 1.0;
        return T;
    }
    for (int mi = 0; mi < ns; ++mi) {
        const int m = mi - l, am = std::abs(m);
        const int v2m = m < 0 ? 1 : 0;
        for (int t = 0; t <= (l - am) / 2; ++t)
            for (int u = 0; u <= t; ++u)
                for (int v2 = v2m; v2 <= am; v2 += 2) {
                    const int sgn = ((t + (v2 - v2m) / 2) & 1) ? -1 : 1;
                    const double c = sgn * std::pow(0.25, t) * binom(l, t) * binom(l - t, am + t) * binom(t, u) *
                                     binom(am, v2);
                    const int px = 2 * t + am - 2 * u - v2, py = 2 * u + v2;
                    const int idx = cart_index(l, px, py);
                    T[(size_t)mi * nc + idx] += c;
                }
    }
    return T;
}

// contracted, same-centre Cartesian overlap of a shell (for normalisation)
static std::vector<double> self_overlap_cart(const Shell& s) {
    const int nc = ncart(s.l);
    std::vector<double> S((size_t)nc * nc, 0.0);
    for (int c1 = 0; c1 < nc; ++c1)
        for (int c2 = 0; c2 < nc; ++c2) {
            int x1, y1, z1, x2, y2, z2;
            cart_exps(s.l, c1, &x1, &y1, &z1);
            cart_exps(s.l, c2, &x2, &y2, &z2);
            double v = 0.0;
            for (int i = 0; i < s.nprim; ++i)
                for (int j = 0; j < s.nprim; ++j) {
                    const double p = s.a[i] + s.a[j];
                    v += s.c[i] * s.c[j] * gauss_moment(x1 + x2, p) * gauss_moment(y1 + y2, p) *
                         gauss_moment(z1 + z2, p);
                }
            S[(size_t)c1 * nc + c2] = v;
        }
    return S;
}

bool finish_shell(Shell& s) {
    for (int i = 0; i < s.nprim; ++i) s.c[i] *= gto_norm(s.l, s.a[i]);
    s.T = cart2sph(s.l);
    const int nc = ncart(s.l), ns = 2 * s.l + 1;
    const std::vector<double> S = self_overlap_cart(s);
    for (int m = 0; m < ns; ++m) {
        double n = 0.0;
        for (int c1 = 0; c1 < nc; ++c1)
            for (int c2 = 0; c2 < nc; ++c2) n += s.T[(size_t)m * nc + c1] * S[(size_t)c1 * nc + c2] * s.T[(size_t)m * nc + c2];
        if (!(n > 0)) return false;
        const double f = 1.0 / std::sqrt(n);
        for (int c = 0; c < nc; ++c) s.T[(size_t)m * nc + c] *= f;
    }
    s.nsph = ns;
    return true;
}

// ---------------------------------------------------------------------------
// one-electron integrals of a shell pair, Cartesian then transformed
// ---------------------------------------------------------------------------
static void cart_to_sph_pair(const Shell& A, const Shell& B, const double* cart, double* sph) {
    const int nca = ncart(A.l), ncb = ncart(B.l);
    std::vector<double> tmp((size_t)A.nsph * ncb, 0.0);
    for (int m = 0; m < A.nsph; ++m)
        for (int cb = 0; cb < ncb; ++cb) {
            double v = 0.0;
            for (int ca = 0; ca < nca; ++ca) v += A.T[(size_t)m * nca + ca] * cart[(size_t)ca * ncb + cb];
            tmp[(size_t)m * ncb + cb] = v;
        }
    for (int m = 0; m < A.nsph; ++m)
        for (int n = 0; n < B.nsph; ++n) {
            double v = 0.0;
            for (int cb = 0; cb < ncb; ++cb) v += tmp[(size_t)m * ncb + cb] * B.T[(size_t)n * ncb + cb];
            sph[(size_t)m * B.nsph + n] = v;
        }
}

void pair_1e(const Shell& A, const Shell& B, size_t ncharge, const double* charge, const double* cxyz, double* S,
             double* T, double* V) {
    const int la = A.l, lb = B.l, nca = ncart(la), ncb = ncart(lb);
    std::vector<double> s((size_t)nca * ncb, 0.0), t((size_t)nca * ncb, 0.0), v((size_t)nca * ncb, 0.0);
    const double AB2 = (A.A[0] - B.A[0]) * (A.A[0] - B.A[0]) + (A.A[1] - B.A[1]) * (A.A[1] - B.A[1]) +
                       (A.A[2] - B.A[2]) * (A.A[2] - B.A[2]);
    HermiteE1 E[3];
    HermiteR R;
    for (int ia = 0; ia < A.nprim; ++ia)
        for (int ib = 0; ib < B.nprim; ++ib) {
            const double a = A.a[ia], b = B.a[ib], p = a + b, mu = a * b / p;
            const double K = A.c[ia] * B.c[ib] * std::exp(-mu * AB2);
            if (std::fabs(K) < kPairCut) continue;
            double P[3];
            for (int d = 0; d < 3; ++d) {
                P[d] = (a * A.A[d] + b * B.A[d]) / p;
                E[d].build(la, lb + 2, p, P[d] - A.A[d], P[d] - B.A[d]);
            }
            const double ov = K * std::pow(M_PI / p, 1.5);
            for (int ca = 0; ca < nca; ++ca) {
                int ax[3];
                cart_exps(la, ca, &ax[0], &ax[1], &ax[2]);
                for (int cb = 0; cb < ncb; ++cb) {
                    int bx[3];
                    cart_exps(lb, cb, &bx[0], &bx[1], &bx[2]);
                    double s1[3], t1[3];
                    for (int d = 0; d < 3; ++d) {
                        const int i = ax[d], j = bx[d];
                        s1[d] = E[d].e[i][j][0];
                        double kin = -2.0 * b * b * E[d].e[i][j + 2][0] + b * (2 * j + 1) * E[d].e[i][j][0];
                        if (j >= 2) kin -= 0.5 * j * (j - 1) * E[d].e[i][j - 2][0];
                        t1[d] = kin;
                    }
                    s[(size_t)ca * ncb + cb] += ov * s1[0] * s1[1] * s1[2];
                    t[(size_t)ca * ncb + cb] +=
                        ov * (t1[0] * s1[1] * s1[2] + s1[0] * t1[1] * s1[2] + s1[0] * s1[1] * t1[2]);
                }
            }
            if (V) {
                const double pref = K * 2.0 * M_PI / p;
                for (size_t q = 0; q < ncharge; ++q) {
                    R.build(la + lb, p, P[0] - cxyz[3 * q], P[1] - cxyz[3 * q + 1], P[2] - cxyz[3 * q + 2]);
                    for (int ca = 0; ca < nca; ++ca) {
                        int ax[3];
                        cart_exps(la, ca, &ax[0], &ax[1], &ax[2]);
                        for (int cb = 0; cb < ncb; ++cb) {
                            int bx[3];
                            cart_exps(lb, cb, &bx[0], &bx[1], &bx[2]);
                            double acc = 0.0;
                            for (int tt = 0; tt <= ax[0] + bx[0]; ++tt)
                                for (int uu = 0; uu <= ax[1] + bx[1]; ++uu)
                                    for (int vv = 0; vv <= ax[2] + bx[2]; ++vv)
                                        acc += E[0].e[ax[0]][bx[0]][tt] * E[1].e[ax[1]][bx[1]][uu] *
                                               E[2].e[ax[2]][bx[2]][vv] * R.get(tt, uu, vv);
                            v[(size_t)ca * ncb + cb] -= charge[q] * pref * acc;
                        }
                    }
                }
            }
        }
    if (S) cart_to_sph_pair(A, B, s.data(), S);
    if (T) cart_to_sph_pair(A, B, t.data(), T);
    if (V) cart_to_sph_pair(A, B, v.data(), V);
}

// ---------------------------------------------------------------------------
// Coulomb-type integrals.  With Hermite expansions E^{ab} (bra) and E^{c}
// (ket, one centre):  (ab|c) = 2π^{5/2}/(p q sqrt(p+q)) Σ_tuv E^{ab}_tuv
//                              Σ_τνφ (-1)^{τ+ν+φ} E^c_τνφ R_{t+τ,u+ν,v+φ}(α, P−C)
// ---------------------------------------------------------------------------

// G[c][t][u][v] = Σ_τνφ (-1)^{τ+ν+φ} E^c_τ E^c_ν E^c_φ R_{t+τ,u+ν,v+φ} for t,u,v <= Lb (bra Hermite range)
static void ket_contract(int lc, const HermiteE0& Ec, const HermiteR& R, int Lb, double* G) {
    const int ncc = ncart(lc), nb = Lb + 1;
    for (int cc = 0; cc < ncc; ++cc) {
        int cx, cy, cz;
        cart_exps(lc, cc, &cx, &cy, &cz);
        double* g = G + (size_t)cc * nb * nb * nb;
        for (int t = 0; t <= Lb; ++t)
            for (int u = 0; u <= Lb - t; ++u)
                for (int v = 0; v <= Lb - t - u; ++v) {
                    double acc = 0.0;
                    for (int ta = cx & 1; ta <= cx; ta += 2) {
                        const double ex = Ec.e[cx][ta];
                        for (int ua = cy & 1; ua <= cy; ua += 2) {
                            const double exy = ex * Ec.e[cy][ua];
                            for (int va = cz & 1; va <= cz; va += 2) {
                                const double sgn = ((ta + ua + va) & 1) ? -1.0 : 1.0;
                                acc += sgn * exy * Ec.e[cz][va] * R.get(t + ta, u + ua, v + va);
                            }
                        }
                    }
                    g[(t * nb + u) * nb + v] = acc;
                }
    }
}

void aux_2c(const Shell& C, const Shell& D, double* out /* nsph_c × nsph_d */) {
    // bra C (centre P = C), ket D (centre Q = D), R at C − D, ket sign on D
    const int lc = C.l, ld = D.l, ncc = ncart(lc), ncd = ncart(ld);
    std::vector<double> cart((size_t)ncc * ncd, 0.0);
    const double X = C.A[0] - D.A[0], Y = C.A[1] - D.A[1], Z = C.A[2] - D.A[2];
    HermiteE0 Ec, Ed;
    HermiteR R;
    const int nb = lc + 1;
    std::vector<double> G((size_t)ncd * nb * nb * nb);
    for (int ic = 0; ic < C.nprim; ++ic)
        for (int id = 0; id < D.nprim; ++id) {
            const double g = C.a[ic], d = D.a[id];
            const double alpha = g * d / (g + d);
            const double pref = 2.0 * std::pow(M_PI, 2.5) / (g * d * std::sqrt(g + d)) * C.c[ic] * D.c[id];
            Ec.build(lc, g);
            Ed.build(ld, d);
            R.build(lc + ld, alpha, X, Y, Z);
            ket_contract(ld, Ed, R, lc, G.data());
            for (int cc = 0; cc < ncc; ++cc) {
                int cx, cy, cz;
                cart_exps(lc, cc, &cx, &cy, &cz);
                for (int cd = 0; cd < ncd; ++cd) {
                    const double* gg = G.data() + (size_t)cd * nb * nb * nb;
                    double acc = 0.0;
                    for (int t = cx & 1; t <= cx; t += 2)
                        for (int u = cy & 1; u <= cy; u += 2)
                            for (int v = cz & 1; v <= cz; v += 2)
                                acc += Ec.e[cx][t] * Ec.e[cy][u] * Ec.e[cz][v] * gg[(t * nb + u) * nb + v];
                    cart[(size_t)cc * ncd + cd] += pref * acc;
                }
            }
        }
    cart_to_sph_pair(C, D, cart.data(), out);
}

// (AB|C) for one shell pair and one auxiliary shell, spherical, layout
// out[(ma * nsph_b + mb) * nsph_c + mc]
void pair_3c(const Shell& A, const Shell& B, const Shell& C, double* out, std::vector<double>& scratch) {
    const int la = A.l, lb = B.l, lc = C.l;
    const int nca = ncart(la), ncb = ncart(lb), ncc = ncart(lc);
    const int Lb = la + lb, nb = Lb + 1;
    const size_t ncart3 = (size_t)nca * ncb * ncc;
    const size_t gsz = (size_t)ncc * nb * nb * nb;
    if (scratch.size() < ncart3 + gsz) scratch.resize(ncart3 + gsz);
    double* cart = scratch.data();
    double* G = cart + ncart3;
    std::fill(cart, cart + ncart3, 0.0);
    const double AB2 = (A.A[0] - B.A[0]) * (A.A[0] - B.A[0]) + (A.A[1] - B.A[1]) * (A.A[1] - B.A[1]) +
                       (A.A[2] - B.A[2]) * (A.A[2] - B.A[2]);
    HermiteE1 E[3];
    HermiteE0 Ec;
    HermiteR R;
    for (int ia = 0; ia < A.nprim; ++ia)
        for (int ib = 0; ib < B.nprim; ++ib) {
            const double a = A.a[ia], b = B.a[ib], p = a + b, mu = a * b / p;
            const double K = A.c[ia] * B.c[ib] * std::exp(-mu * AB2);
            if (std::fabs(K) < kPairCut) continue;
            double P[3];
            for (int d = 0; d < 3; ++d) {
                P[d] = (a * A.A[d] + b * B.A[d]) / p;
                E[d].build(la, lb, p, P[d] - A.A[d], P[d] - B.A[d]);
            }
            for (int ic = 0; ic < C.nprim; ++ic) {
                const double g = C.a[ic];
                const double alpha = p * g / (p + g);
                const double pref = 2.0 * std::pow(M_PI, 2.5) / (p * g * std::sqrt(p + g)) * K * C.c[ic];
                Ec.build(lc, g);
                R.build(Lb + lc, alpha, P[0] - C.A[0], P[1] - C.A[1], P[2] - C.A[2]);
                ket_contract(lc, Ec, R, Lb, G);
                for (int ca = 0; ca < nca; ++ca) {
                    int ax, ay, az;
                    cart_exps(la, ca, &ax, &ay, &az);
                    for (int cb = 0; cb < ncb; ++cb) {
                        int bx, by, bz;
                        cart_exps(lb, cb, &bx, &by, &bz);
                        double* o = cart + ((size_t)ca * ncb + cb) * ncc;
                        for (int t = 0; t <= ax + bx; ++t) {
                            const double et = E[0].e[ax][bx][t];
                            for (int u = 0; u <= ay + by; ++u) {
                                const double etu = et * E[1].e[ay][by][u];
                                for (int v = 0; v <= az + bz; ++v) {
                                    const double e3 = pref * etu * E[2].e[az][bz][v];
                                    const size_t gi = (size_t)(t * nb + u) * nb + v;
                                    for (int cc = 0; cc < ncc; ++cc) o[cc] += e3 * G[(size_t)cc * nb * nb * nb + gi];
                                }
                            }
                        }
                    }
                }
            }
        }
    // transform a, b, c to spherical
    const int nsa = A.nsph, nsb = B.nsph, nsc = C.nsph;
    std::vector<double> t1((size_t)nca * ncb * nsc, 0.0), t2((size_t)nca * nsb * nsc, 0.0);
    for (size_t ab = 0; ab < (size_t)nca * ncb; ++ab)
        for (int mc = 0; mc < nsc; ++mc) {
            double v = 0.0;
            for (int cc = 0; cc < ncc; ++cc) v += C.T[(size_t)mc * ncc + cc] * cart[ab * ncc + cc];
            t1[ab * nsc + mc] = v;
        }
    for (int ca = 0; ca < nca; ++ca)
        for (int mb = 0; mb < nsb; ++mb)
            for (int mc = 0; mc < nsc; ++mc) {
                double v = 0.0;
                for (int cb = 0; cb < ncb; ++cb)
                    v += B.T[(size_t)mb * ncb + cb] * t1[((size_t)ca * ncb + cb) * nsc + mc];
                t2[((size_t)ca * nsb + mb) * nsc + mc] = v;
            }
    for (int ma = 0; ma < nsa; ++ma)
        for (int mb = 0; mb < nsb; ++mb)
            for (int mc = 0; mc < nsc; ++mc) {
                double v = 0.0;
                for (int ca = 0; ca < nca; ++ca)
                    v += A.T[(size_t)ma * nca + ca] * t2[((size_t)ca * nsb + mb) * nsc + mc];
                out[((size_t)ma * nsb + mb) * nsc + mc] = v;
            }
}

// Schwarz-type bound for a shell pair: max |c_a c_b| exp(-μ R²) over its primitive pairs
double pair_bound(const Shell& A, const Shell& B) {
    const double AB2 = (A.A[0] - B.A[0]) * (A.A[0] - B.A[0]) + (A.A[1] - B.A[1]) * (A.A[1] - B.A[1]) +
                       (A.A[2] - B.A[2]) * (A.A[2] - B.A[2]);
    double m = 0.0;
    for (int ia = 0; ia < A.nprim; ++ia)
        for (int ib = 0; ib < B.nprim; ++ib) {
            const double mu = A.a[ia] * B.a[ib] / (A.a[ia] + B.a[ib]);
            m = std::max(m, std::fabs(A.c[ia] * B.c[ib]) * std::exp(-mu * AB2));
        }
    return m;
}

void int3c_slab(const Basis& orb, const Basis& aux, size_t j0, size_t j1, double* out) {
    const size_t q0 = j0 * (j0 + 1) / 2, q1 = j1 * (j1 + 1) / 2, nq = q1 - q0;
    const size_t naux = aux.nbf;
    std::memset(out, 0, sizeof(double) * nq * naux);
    // shell pairs (sa <= sb) with sb's functions inside [j0, j1)
    std::vector<std::pair<int, int>> pairs;
    for (int sb = 0; sb < (int)orb.sh.size(); ++sb) {
        const Shell& B = orb.sh[sb];
        if ((size_t)B.off < j0 || (size_t)B.off >= j1) continue;
        for (int sa = 0; sa <= sb; ++sa)
            if (pair_bound(orb.sh[sa], B) >= kPairCut) pairs.emplace_back(sa, sb);
    }
#pragma omp parallel
    {
        std::vector<double> scratch, blk;
#pragma omp for schedule(dynamic, 1)
        for (long long ip = 0; ip < (long long)pairs.size(); ++ip) {
            const Shell& A = orb.sh[pairs[ip].first];
            const Shell& B = orb.sh[pairs[ip].second];
            for (const Shell& Cs : aux.sh) {
                blk.resize((size_t)A.nsph * B.nsph * Cs.nsph);
                pair_3c(A, B, Cs, blk.data(), scratch);
                for (int ma = 0; ma < A.nsph; ++ma)
                    for (int mb = 0; mb < B.nsph; ++mb) {
                        const size_t i = A.off + ma, j = B.off + mb;
                        if (i > j) continue;   // same shell: upper triangle only
                        const size_t q = j * (j + 1) / 2 + i - q0;
                        for (int mc = 0; mc < Cs.nsph; ++mc)
                            out[(size_t)(Cs.off + mc) * nq + q] = blk[((size_t)ma * B.nsph + mb) * Cs.nsph + mc];
                    }
            }
        }
    }
}

}  // namespace rsi

using namespace rsi;

extern "C" const char* rsi_last_error(void) { return last_error().c_str(); }

extern "C" int rsi_cart2sph(int l, double* out) {
    if (l < 0 || l > kMaxL || !out) {
        set_error("rsi_cart2sph: l must be 0..4 and out non-NULL");
        return RJK_ERR_ARG;
    }
    const std::vector<double> T = cart2sph(l);
    std::copy(T.begin(), T.end(), out);
    return RJK_OK;
}

extern "C" int rsi_basis_create(size_t nshell, const int* shell_atom, const int* shell_l, const int* shell_nprim,
                                const double* exps, const double* coefs, size_t natm, const double* atom_xyz,
                                rsi_basis** out) {
    if (!out || !shell_atom || !shell_l || !shell_nprim || !exps || !coefs || !atom_xyz) {
        set_error("rsi_basis_create: NULL argument");
        return RJK_ERR_ARG;
    }
    *out = nullptr;
    rsi_basis* b = new rsi_basis();
    size_t off = 0, poff = 0;
    for (size_t s = 0; s < nshell; ++s) {
        Shell sh;
        sh.l = shell_l[s];
        sh.nprim = shell_nprim[s];
        sh.atom = shell_atom[s];
        if (sh.l < 0 || sh.l > kMaxL || sh.nprim <= 0 || sh.atom < 0 || (size_t)sh.atom >= natm) {
            delete b;
            set_error("rsi_basis_create: shell " + std::to_string(s) + " has an invalid l, nprim or atom");
            return RJK_ERR_ARG;
        }
        for (int d = 0; d < 3; ++d) sh.A[d] = atom_xyz[3 * sh.atom + d];
        sh.a.assign(exps + poff, exps + poff + sh.nprim);
        sh.c.assign(coefs + poff, coefs + poff + sh.nprim);
        poff += sh.nprim;
        for (double a : sh.a)
            if (!(a > 0) || !std::isfinite(a)) {
                delete b;
                set_error("rsi_basis_create: non-positive exponent in shell " + std::to_string(s));
                return RJK_ERR_ARG;
            }
        if (!finish_shell(sh)) {
            delete b;
            set_error("rsi_basis_create: shell " + std::to_string(s) + " has zero norm");
            return RJK_ERR_ARG;
        }
        sh.off = (int)off;
        off += sh.nsph;
        b->sh.push_back(std::move(sh));
    }
    b->nbf = off;
    *out = b;
    return RJK_OK;
}

extern "C" int rsi_basis_destroy(rsi_basis* b) {
    delete b;
    return RJK_OK;
}

extern "C" size_t rsi_basis_nbf(const rsi_basis* b) { return b ? b->nbf : 0; }

extern "C" int rsi_int1e(const rsi_basis* orb, size_t ncharge, const double* charge, const double* charge_xyz,
                         double* S, double* T, double* V) {
    if (!orb) {
        set_error("rsi_int1e: NULL basis");
        return RJK_ERR_ARG;
    }
    const size_t n = orb->nbf;
    const int ns = (int)orb->sh.size();
    std::vector<std::pair<int, int>> pairs;
    for (int b = 0; b < ns; ++b)
        for (int a = 0; a <= b; ++a) pairs.emplace_back(a, b);
#pragma omp parallel
    {
        std::vector<double> s, t, v;
#pragma omp for schedule(dynamic, 4)
        for (long long ip = 0; ip < (long long)pairs.size(); ++ip) {
            const Shell& A = orb->sh[pairs[ip].first];
            const Shell& B = orb->sh[pairs[ip].second];
            const size_t m = (size_t)A.nsph * B.nsph;
            s.assign(m, 0.0);
            t.assign(m, 0.0);
            v.assign(m, 0.0);
            pair_1e(A, B, ncharge, charge, charge_xyz, s.data(), t.data(), V ? v.data() : nullptr);
            for (int ma = 0; ma < A.nsph; ++ma)
                for (int mb = 0; mb < B.nsph; ++mb) {
                    const size_t i = A.off + ma, j = B.off + mb, k = (size_t)ma * B.nsph + mb;
                    if (S) S[i + j * n] = S[j + i * n] = s[k];
                    if (T) T[i + j * n] = T[j + i * n] = t[k];
                    if (V) V[i + j * n] = V[j + i * n] = v[k];
                }
        }
    }
    return RJK_OK;
}

extern "C" int rsi_int2c2e(const rsi_basis* aux, double* V) {
    if (!aux || !V) {
        set_error("rsi_int2c2e: NULL argument");
        return RJK_ERR_ARG;
    }
    const size_t n = aux->nbf;
    const int ns = (int)aux->sh.size();
    std::vector<std::pair<int, int>> pairs;
    for (int b = 0; b < ns; ++b)
        for (int a = 0; a <= b; ++a) pairs.emplace_back(a, b);
#pragma omp parallel
    {
        std::vector<double> blk;
#pragma omp for schedule(dynamic, 8)
        for (long long ip = 0; ip < (long long)pairs.size(); ++ip) {
            const Shell& C = aux->sh[pairs[ip].first];
            const Shell& D = aux->sh[pairs[ip].second];
            blk.assign((size_t)C.nsph * D.nsph, 0.0);
            aux_2c(C, D, blk.data());
            for (int mc = 0; mc < C.nsph; ++mc)
                for (int md = 0; md < D.nsph; ++md) {
                    const size_t i = C.off + mc, j = D.off + md;
                    V[i + j * n] = V[j + i * n] = blk[(size_t)mc * D.nsph + md];
                }
        }
    }
    return RJK_OK;
}

extern "C" int rsi_int3c2e_packed(const rsi_basis* orb, const rsi_basis* aux, size_t j0, size_t j1, double* out) {
    if (!orb || !aux || !out) {
        set_error("rsi_int3c2e_packed: NULL argument");
        return RJK_ERR_ARG;
    }
    if (j0 > j1 || j1 > orb->nbf || !on_shell_boundary(*orb, j0) || !on_shell_boundary(*orb, j1)) {
        set_error("rsi_int3c2e_packed: [j0, j1) must lie on shell boundaries of the orbital basis");
        return RJK_ERR_ARG;
    }
    int3c_slab(*orb, *aux, j0, j1, out);
    return RJK_OK;
}
