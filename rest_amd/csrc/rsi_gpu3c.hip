// rsi_gpu3c.hip — three-centre integrals (μν|P) on the GPU for the B producer.
//
// One thread per (orbital shell pair, auxiliary shell); kernels are
// instantiated per angular-momentum class (l_a >= l_b, l_c) so every loop has
// compile-time bounds.  McMurchie–Davidson as in rsi_core.hpp: Hermite E
// coefficients of the pair, single-centre E of the auxiliary Gaussian,
// Hermite integrals R_tuv from the Boys function (tabulated, 7-term Taylor),
// computed in place with one array (level n overwrites level n+1 in order of
// decreasing t+u+v).  Results are transformed to spherical functions and
// written into the pair slab X[P][q] that feeds the V^{-1/2} GEMM.
#include <algorithm>
#include <cmath>
#include <utility>
#include <vector>

#include "rjk_internal.hpp"
#include "rsi_core.hpp"
#include "rsi_gpu3c.hpp"
#include "rsi_internal.hpp"

namespace rsi {
namespace {

__host__ __device__ constexpr int nc_(int l) { return (l + 1) * (l + 2) / 2; }

__device__ inline void cexp(int l, int c, int* x, int* y, int* z) {
    int idx = 0;
    for (int a = l; a >= 0; --a)
        for (int b = l - a; b >= 0; --b) {
            if (idx == c) {
                *x = a;
                *y = b;
                *z = l - a - b;
                return;
            }
            ++idx;
        }
}

// Boys F_0..F_M(T) from the device table (same grid and expansion as rsi::Boys)
template <int M>
__device__ inline void boys_dev(const double* __restrict__ tab, double T, double* F) {
    if (T < Boys::kTMax) {
        const int i = (int)(T / Boys::kStep + 0.5);
        const double dt = i * Boys::kStep - T;
        const double* row = tab + (size_t)i * Boys::kMTab;
#pragma unroll
        for (int m = 0; m <= M; ++m) {
            double s = 0.0, p = 1.0;
#pragma unroll
            for (int k = 0; k <= Boys::kOrd; ++k) {
                s += row[m + k] * p;
                p *= dt / (k + 1);
            }
            F[m] = s;
        }
    } else {
        const double e = exp(-T);
        F[0] = 0.5 * sqrt(M_PI / T);
#pragma unroll
        for (int m = 0; m < M; ++m) F[m + 1] = ((2 * m + 1) * F[m] - e) / (2.0 * T);
    }
}

__host__ __device__ constexpr int rbase(int s) { return s * (s + 1) * (s + 2) / 6; }
__host__ __device__ constexpr int ridx(int t, int u, int v) {
    return rbase(t + u + v) + (u + v) * (u + v + 1) / 2 + v;
}

// R_tuv(α, X, Y, Z) for t+u+v <= L, in place (index ridx)
template <int L>
__device__ inline void hermite_r(const double* __restrict__ boys, double alpha, double X, double Y, double Z,
                                 double* r) {
    double F[L + 1], pw[L + 1];
    boys_dev<L>(boys, alpha * (X * X + Y * Y + Z * Z), F);
    pw[0] = 1.0;
    for (int n = 1; n <= L; ++n) pw[n] = pw[n - 1] * (-2.0 * alpha);
    r[0] = pw[L] * F[L];
    for (int n = L - 1; n >= 0; --n) {
        // level n from level n+1: sums s = L-n .. 1 in decreasing order (each reads only
        // sums s-1 and s-2, still at level n+1), then s = 0
        for (int s = L - n; s >= 1; --s)
            for (int t = s; t >= 0; --t)
                for (int u = s - t; u >= 0; --u) {
                    const int v = s - t - u;
                    double val;
                    if (t > 0) {
                        val = X * r[ridx(t - 1, u, v)];
                        if (t > 1) val += (t - 1) * r[ridx(t - 2, u, v)];
                    } else if (u > 0) {
                        val = Y * r[ridx(t, u - 1, v)];
                        if (u > 1) val += (u - 1) * r[ridx(t, u - 2, v)];
                    } else {
                        val = Z * r[ridx(t, u, v - 1)];
                        if (v > 1) val += (v - 1) * r[ridx(t, u, v - 2)];
                    }
                    r[ridx(t, u, v)] = val;
                }
        r[0] = pw[n] * F[n];
    }
}

// 1-D Hermite coefficients of a pair, E[i][j][t]
template <int LA, int LB>
__device__ inline void hermite_e(double p, double XPA, double XPB, double (*E)[LB + 1][LA + LB + 1]) {
    const double oo2p = 0.5 / p;
    for (int i = 0; i <= LA; ++i)
        for (int j = 0; j <= LB; ++j)
            for (int t = 0; t <= LA + LB; ++t) E[i][j][t] = 0.0;
    E[0][0][0] = 1.0;
    for (int i = 0; i <= LA; ++i)
        for (int j = 0; j <= LB; ++j) {
            if (i == 0 && j == 0) continue;
            const int pi = j == 0 ? i - 1 : i, pj = j == 0 ? j : j - 1;
            const double X = j == 0 ? XPA : XPB;
            for (int t = 0; t <= i + j; ++t) {
                double v = X * (t <= pi + pj ? E[pi][pj][t] : 0.0);
                if (t > 0) v += oo2p * E[pi][pj][t - 1];
                if (t + 1 <= pi + pj) v += (t + 1) * E[pi][pj][t + 1];
                E[i][j][t] = v;
            }
        }
}

template <int LA, int LB, int LC>
__global__ __launch_bounds__(128) void k_int3c(const GShell* __restrict__ orb, const GShell* __restrict__ aux,
                                               const double* __restrict__ pexp, const double* __restrict__ pcoef,
                                               const double* __restrict__ tsph, const int2* __restrict__ pairs,
                                               int npairs, const int* __restrict__ auxl, int nauxl,
                                               const double* __restrict__ boys, size_t q0, size_t nq,
                                               double* __restrict__ X) {
    constexpr int NA = nc_(LA), NB = nc_(LB), NC = nc_(LC), LAB = LA + LB, L = LAB + LC;
    constexpr int NR = (L + 1) * (L + 2) * (L + 3) / 6;
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (size_t)npairs * nauxl) return;
    const int ip = (int)(t % npairs), icx = (int)(t / npairs);
    const int2 pr = pairs[ip];
    const GShell& A = orb[pr.x];
    const GShell& B = orb[pr.y];
    const GShell& C = aux[auxl[icx]];
    double acc[NA * NB * NC];
#pragma unroll
    for (int k = 0; k < NA * NB * NC; ++k) acc[k] = 0.0;
    const double ABx = A.A[0] - B.A[0], ABy = A.A[1] - B.A[1], ABz = A.A[2] - B.A[2];
    const double AB2 = ABx * ABx + ABy * ABy + ABz * ABz;
    double Ex[LA + 1][LB + 1][LAB + 1], Ey[LA + 1][LB + 1][LAB + 1], Ez[LA + 1][LB + 1][LAB + 1];
    double Ec[LC + 1][LC + 1];
    double r[NR];
    for (int ia = 0; ia < A.nprim; ++ia)
        for (int ib = 0; ib < B.nprim; ++ib) {
            const double a = pexp[A.p0 + ia], b = pexp[B.p0 + ib], p = a + b, mu = a * b / p;
            const double K = pcoef[A.p0 + ia] * pcoef[B.p0 + ib] * exp(-mu * AB2);
            if (fabs(K) < kPairCut) continue;
            const double Px = (a * A.A[0] + b * B.A[0]) / p, Py = (a * A.A[1] + b * B.A[1]) / p,
                         Pz = (a * A.A[2] + b * B.A[2]) / p;
            hermite_e<LA, LB>(p, Px - A.A[0], Px - B.A[0], Ex);
            hermite_e<LA, LB>(p, Py - A.A[1], Py - B.A[1], Ey);
            hermite_e<LA, LB>(p, Pz - A.A[2], Pz - B.A[2], Ez);
            for (int ic = 0; ic < C.nprim; ++ic) {
                const double g = pexp[C.p0 + ic];
                const double alpha = p * g / (p + g);
                const double pref = 2.0 * 17.493418327624862 / (p * g * sqrt(p + g)) * K * pcoef[C.p0 + ic];  // 2π^{5/2}
                {   // single-centre Hermite coefficients of the auxiliary Gaussian
                    const double oo2g = 0.5 / g;
                    for (int k = 0; k <= LC; ++k)
                        for (int tt = 0; tt <= LC; ++tt) Ec[k][tt] = 0.0;
                    Ec[0][0] = 1.0;
                    for (int k = 1; k <= LC; ++k)
                        for (int tt = 0; tt <= k; ++tt) {
                            double v = 0.0;
                            if (tt > 0) v += oo2g * Ec[k - 1][tt - 1];
                            if (tt + 1 <= k - 1) v += (tt + 1) * Ec[k - 1][tt + 1];
                            Ec[k][tt] = v;
                        }
                }
                hermite_r<L>(boys, alpha, Px - C.A[0], Py - C.A[1], Pz - C.A[2], r);
                for (int cc = 0; cc < NC; ++cc) {
                    int cx, cy, cz;
                    cexp(LC, cc, &cx, &cy, &cz);
                    for (int ca = 0; ca < NA; ++ca) {
                        int ax, ay, az;
                        cexp(LA, ca, &ax, &ay, &az);
                        for (int cb = 0; cb < NB; ++cb) {
                            int bx, by, bz;
                            cexp(LB, cb, &bx, &by, &bz);
                            double sum = 0.0;
                            for (int tx = 0; tx <= ax + bx; ++tx)
                                for (int uy = 0; uy <= ay + by; ++uy) {
                                    const double exy = Ex[ax][bx][tx] * Ey[ay][by][uy];
                                    for (int vz = 0; vz <= az + bz; ++vz) {
                                        const double e3 = exy * Ez[az][bz][vz];
                                        double g3 = 0.0;
                                        for (int ta = cx & 1; ta <= cx; ta += 2)
                                            for (int ua = cy & 1; ua <= cy; ua += 2)
                                                for (int va = cz & 1; va <= cz; va += 2) {
                                                    const double sgn = ((ta + ua + va) & 1) ? -1.0 : 1.0;
                                                    g3 += sgn * Ec[cx][ta] * Ec[cy][ua] * Ec[cz][va] *
                                                          r[ridx(tx + ta, uy + ua, vz + va)];
                                                }
                                        sum += e3 * g3;
                                    }
                                }
                            acc[(ca * NB + cb) * NC + cc] += pref * sum;
                        }
                    }
                }
            }
        }
    // spherical transform and scatter into the slab
    constexpr int SA = 2 * LA + 1, SB = 2 * LB + 1, SC = 2 * LC + 1;
    const double* __restrict__ Ta = tsph + A.t0;
    const double* __restrict__ Tb = tsph + B.t0;
    const double* __restrict__ Tc = tsph + C.t0;
    const bool same = pr.x == pr.y;
    for (int ma = 0; ma < SA; ++ma)
        for (int mb = 0; mb < SB; ++mb) {
            const size_t i = A.off + ma, j = B.off + mb;
            if (same && i > j) continue;
            const size_t q = (i <= j ? j * (j + 1) / 2 + i : i * (i + 1) / 2 + j) - q0;
            for (int mc = 0; mc < SC; ++mc) {
                double v = 0.0;
                for (int ca = 0; ca < NA; ++ca) {
                    const double ta = Ta[ma * NA + ca];
                    if (ta == 0.0) continue;
                    for (int cb = 0; cb < NB; ++cb) {
                        const double tb = Tb[mb * NB + cb];
                        if (tb == 0.0) continue;
                        double s = 0.0;
                        for (int cc = 0; cc < NC; ++cc) s += Tc[mc * NC + cc] * acc[(ca * NB + cb) * NC + cc];
                        v += ta * tb * s;
                    }
                }
                X[(size_t)(C.off + mc) * nq + q] = v;
            }
        }
}

template <int LA, int LB, int LC>
hipError_t launch_class(hipStream_t st, const Gpu3c& g, const int2* pairs, int npairs, const int* auxl, int nauxl,
                        size_t q0, size_t nq, double* X) {
    const size_t n = (size_t)npairs * nauxl;
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL((k_int3c<LA, LB, LC>), dim3((unsigned)((n + 127) / 128)), dim3(128), 0, st, g.orb, g.aux,
                       g.pexp, g.pcoef, g.tsph, pairs, npairs, auxl, nauxl, g.boys, q0, nq, X);
    return hipGetLastError();
}

template <int LA, int LB>
hipError_t launch_lc(hipStream_t st, const Gpu3c& g, int lc, const int2* pairs, int npairs, const int* auxl,
                     int nauxl, size_t q0, size_t nq, double* X) {
    switch (lc) {
        case 0: return launch_class<LA, LB, 0>(st, g, pairs, npairs, auxl, nauxl, q0, nq, X);
        case 1: return launch_class<LA, LB, 1>(st, g, pairs, npairs, auxl, nauxl, q0, nq, X);
        case 2: return launch_class<LA, LB, 2>(st, g, pairs, npairs, auxl, nauxl, q0, nq, X);
        case 3: return launch_class<LA, LB, 3>(st, g, pairs, npairs, auxl, nauxl, q0, nq, X);
        case 4: return launch_class<LA, LB, 4>(st, g, pairs, npairs, auxl, nauxl, q0, nq, X);
    }
    return hipErrorInvalidValue;
}

}  // namespace

bool gpu3c_supported(const Basis& orb, const Basis& aux) {
    for (const Shell& s : orb.sh)
        if (s.l > kGpuMaxLOrb) return false;
    for (const Shell& s : aux.sh)
        if (s.l > kGpuMaxLAux) return false;
    return true;
}

int gpu3c_init(Gpu3c* g, const Basis& orb, const Basis& aux, const std::vector<size_t>& cuts, hipStream_t st) {
    std::vector<GShell> go, ga;
    std::vector<double> ex, co, ts;
    auto add = [&](const Shell& s, std::vector<GShell>& out) {
        GShell q{};
        for (int d = 0; d < 3; ++d) q.A[d] = s.A[d];
        q.l = s.l;
        q.nprim = s.nprim;
        q.p0 = (int)ex.size();
        q.t0 = (int)ts.size();
        q.off = s.off;
        ex.insert(ex.end(), s.a.begin(), s.a.end());
        co.insert(co.end(), s.c.begin(), s.c.end());
        ts.insert(ts.end(), s.T.begin(), s.T.end());
        out.push_back(q);
    };
    for (const Shell& s : orb.sh) add(s, go);
    for (const Shell& s : aux.sh) add(s, ga);
    const size_t nrow = (size_t)(Boys::kTMax / Boys::kStep) + 2;
    std::vector<double> tab(nrow * Boys::kMTab);
    for (size_t i = 0; i < nrow; ++i) Boys::series(i * Boys::kStep, Boys::kMTab - 1, &tab[i * Boys::kMTab]);
    // auxiliary shells grouped by l
    std::vector<int> flat;
    for (int l = 0; l <= kGpuMaxLAux; ++l) {
        g->aux_off[l] = (int)flat.size();
        for (size_t i = 0; i < aux.sh.size(); ++i)
            if (aux.sh[i].l == l) flat.push_back((int)i);
        g->aux_cnt[l] = (int)flat.size() - g->aux_off[l];
    }
    // screened shell pairs per slab and class; A is the shell with the larger l
    g->cuts = cuts;
    const size_t nslab = cuts.size() - 1;
    g->cls_off.assign(nslab * kGpuClasses, 0);
    g->cls_cnt.assign(nslab * kGpuClasses, 0);
    std::vector<int2> all;
    for (size_t k = 0; k < nslab; ++k) {
        std::vector<int2> cls[kGpuClasses];
        for (int sb = 0; sb < (int)orb.sh.size(); ++sb) {
            const Shell& B = orb.sh[sb];
            if ((size_t)B.off < cuts[k] || (size_t)B.off >= cuts[k + 1]) continue;
            for (int sa = 0; sa <= sb; ++sa) {
                const Shell& A = orb.sh[sa];
                if (pair_bound(A, B) < kPairCut) continue;
                const int la = std::max(A.l, B.l), lb = std::min(A.l, B.l);
                cls[la * (la + 1) / 2 + lb].push_back(A.l >= B.l ? int2{sa, sb} : int2{sb, sa});
            }
        }
        for (int c = 0; c < kGpuClasses; ++c) {
            g->cls_off[k * kGpuClasses + c] = all.size();
            g->cls_cnt[k * kGpuClasses + c] = cls[c].size();
            all.insert(all.end(), cls[c].begin(), cls[c].end());
        }
    }
    auto up = [&](void** dst, const void* src, size_t bytes) -> bool {
        if (hipMalloc(dst, bytes ? bytes : 8) != hipSuccess) return false;
        return !bytes || hipMemcpyAsync(*dst, src, bytes, hipMemcpyHostToDevice, st) == hipSuccess;
    };
    if (!up((void**)&g->orb, go.data(), go.size() * sizeof(GShell)) ||
        !up((void**)&g->aux, ga.data(), ga.size() * sizeof(GShell)) ||
        !up((void**)&g->pexp, ex.data(), ex.size() * sizeof(double)) ||
        !up((void**)&g->pcoef, co.data(), co.size() * sizeof(double)) ||
        !up((void**)&g->tsph, ts.data(), ts.size() * sizeof(double)) ||
        !up((void**)&g->boys, tab.data(), tab.size() * sizeof(double)) ||
        !up((void**)&g->auxl, flat.data(), flat.size() * sizeof(int)) ||
        !up((void**)&g->pairs, all.data(), all.size() * sizeof(int2)))
        return RJK_ERR_NOMEM;
    // the host vectors above die with this call: finish the uploads first
    return hipStreamSynchronize(st) == hipSuccess ? RJK_OK : RJK_ERR_HIP;
}

void gpu3c_free(Gpu3c* g) {
    for (void* p : {(void*)g->orb, (void*)g->aux, (void*)g->pexp, (void*)g->pcoef, (void*)g->tsph, (void*)g->boys,
                    (void*)g->auxl, (void*)g->pairs})
        if (p) (void)hipFree(p);
    *g = Gpu3c();
}

int gpu3c_slab(const Gpu3c& g, size_t k, hipStream_t st, double* X, size_t naux) {
    const size_t j0 = g.cuts[k], j1 = g.cuts[k + 1];
    const size_t q0 = j0 * (j0 + 1) / 2, nq = j1 * (j1 + 1) / 2 - q0;
    if (hipMemsetAsync(X, 0, nq * naux * sizeof(double), st) != hipSuccess) return RJK_ERR_HIP;
    for (int a = 0; a <= kGpuMaxLOrb; ++a)
        for (int b = 0; b <= a; ++b) {
            const int c = a * (a + 1) / 2 + b;
            const int np = (int)g.cls_cnt[k * kGpuClasses + c];
            if (!np) continue;
            const int2* pp = g.pairs + g.cls_off[k * kGpuClasses + c];
            for (int lc = 0; lc <= kGpuMaxLAux; ++lc) {
                const int na = g.aux_cnt[lc];
                if (!na) continue;
                const int* al = g.auxl + g.aux_off[lc];
                hipError_t e;
                switch (c) {
                    case 0: e = launch_lc<0, 0>(st, g, lc, pp, np, al, na, q0, nq, X); break;
                    case 1: e = launch_lc<1, 0>(st, g, lc, pp, np, al, na, q0, nq, X); break;
                    case 2: e = launch_lc<1, 1>(st, g, lc, pp, np, al, na, q0, nq, X); break;
                    case 3: e = launch_lc<2, 0>(st, g, lc, pp, np, al, na, q0, nq, X); break;
                    case 4: e = launch_lc<2, 1>(st, g, lc, pp, np, al, na, q0, nq, X); break;
                    default: e = launch_lc<2, 2>(st, g, lc, pp, np, al, na, q0, nq, X); break;
                }
                if (e != hipSuccess) return RJK_ERR_HIP;
            }
        }
    return RJK_OK;
}

}  // namespace rsi
