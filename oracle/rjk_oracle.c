/*
 * rjk_oracle.c — CPU FP64 restatement of REST's RI-V J/K builders.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker the HIP library is
 * compared against; only tests/, __graft_entry__.smoke() and bench.py's CPU
 * baseline leg may load it.  The product (rest_amd/) never links or calls it.
 *
 * Every builder follows one function of the reference (file:line given at
 * each).  Sums over the pair index and over the auxiliary index use Neumaier
 * compensation so that the oracle is accurate to about one ulp whatever the
 * order the GPU reduces in.
 *
 * Parity pins (tests/test_oracle.py): numpy brute force on unpacked B,
 * naux=1 and D=I closed forms, one-hot B, an asymmetric D that fixes which
 * triangle the packed J reads, linearity, the occupied/density-only cross-route
 * identity, symmetry and positivity of K.
 *
 * Layouts: packed B "rimatr" [naux][npair], p = j(j+1)/2 + i (i <= j)
 * (src/molecule_io/mod.rs:1679-1680); full B "ri3fn" [naux][nbf][nbf] with
 * (i,j,P) at i + nbf*j + nbf*nbf*P (src/scf_io/mod.rs:2400-2408); matrices
 * column-major.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif
#include <dlfcn.h>

/* ------------------------------------------------------------------------ */
/* compensated accumulation                                                 */
/* ------------------------------------------------------------------------ */
typedef struct { double s, c; } nsum;
static inline void nadd(nsum* a, double x) {
    double t = a->s + x;
    if (fabs(a->s) >= fabs(x)) a->c += (a->s - t) + x;
    else a->c += (x - t) + a->s;
    a->s = t;
}
static inline double nval(const nsum* a) { return a->s + a->c; }

static inline size_t pidx(size_t i, size_t j) { /* i <= j */
    return j * (j + 1) / 2 + i;
}

/* ------------------------------------------------------------------------ */
/* seeded generators (shared definition with the HIP synthetic fill)        */
/* ------------------------------------------------------------------------ */
static inline uint64_t splitmix64_next(uint64_t* x) {
    uint64_t z = (*x += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}
static inline uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

typedef struct { uint64_t s[4]; } xoshiro;
static inline uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
static void xo_seed(xoshiro* r, uint64_t seed) {
    uint64_t x = seed;
    for (int i = 0; i < 4; ++i) r->s[i] = splitmix64_next(&x);
}
static uint64_t xo_next(xoshiro* r) {
    uint64_t* s = r->s;
    uint64_t result = rotl(s[1] * 5, 7) * 9;
    uint64_t t = s[1] << 17;
    s[2] ^= s[0]; s[3] ^= s[1]; s[1] ^= s[2]; s[0] ^= s[3];
    s[2] ^= t; s[3] = rotl(s[3], 45);
    return result;
}
static double xo_uniform(xoshiro* r) { return (double)(xo_next(r) >> 11) * 0x1.0p-53; }
static double xo_normal(xoshiro* r) {
    double u1 = xo_uniform(r), u2 = xo_uniform(r);
    return sqrt(-2.0 * log(1.0 - u1)) * cos(2.0 * M_PI * u2);
}

/* Synthetic packed B value at flat index (P*npair + p).  Identical to the
 * device fill of rjk_fill_rimatr_synthetic. */
double orc_synth_b(uint64_t seed, uint64_t flat, uint64_t naux) {
    uint64_t z = mix64(seed + (flat + 1) * 0x9E3779B97F4A7C15ULL);
    double u = (double)(z >> 11) * 0x1.0p-53;
    return (2.0 * u - 1.0) * sqrt(3.0 / (double)naux);
}
void orc_fill_synth_b(uint64_t seed, size_t npair, size_t naux, size_t p_beg, size_t p_cnt,
                      double* out /* [naux_cnt = p_cnt][npair] for P in [p_beg, p_beg+p_cnt) */) {
#pragma omp parallel for schedule(static)
    for (long long P = 0; P < (long long)p_cnt; ++P)
        for (size_t p = 0; p < npair; ++p)
            out[(size_t)P * npair + p] = orc_synth_b(seed, ((size_t)P + p_beg) * npair + p, naux);
}

/* Water-cluster generator.  n waters on the first n sites (x fastest) of a
 * k×k×k cubic lattice, k = smallest with k^3 >= n, spacing 3.1 Å; O jittered
 * uniformly by ±0.2 Å per axis; each monomer (O at origin, H at (±0.7572,
 * 0.5865, 0) Å) rotated about O by a uniform random unit quaternion
 * (Shoemake); the whole cluster is redrawn if any O–O distance < 2.5 Å.
 * Draw order per water: 3 jitter uniforms, then 3 quaternion uniforms.
 * Output xyz[9n] in Å, atoms ordered O,H,H per water; returns attempts. */
int orc_water_cluster(uint64_t seed, int n, double oh_jitter, double hoh_jitter_deg, double* xyz) {
    xoshiro r;
    xo_seed(&r, seed);
    int k = 1;
    while (k * k * k < n) ++k;
    const double spacing = 3.1, jit = 0.2;
    for (int attempt = 1; attempt < 1000; ++attempt) {
        for (int w = 0; w < n; ++w) {
            int ix = w % k, iy = (w / k) % k, iz = w / (k * k);
            double o[3] = {ix * spacing + (2 * xo_uniform(&r) - 1) * jit,
                           iy * spacing + (2 * xo_uniform(&r) - 1) * jit,
                           iz * spacing + (2 * xo_uniform(&r) - 1) * jit};
            double u1 = xo_uniform(&r), u2 = xo_uniform(&r), u3 = xo_uniform(&r);
            double qa = sqrt(1 - u1) * sin(2 * M_PI * u2), qb = sqrt(1 - u1) * cos(2 * M_PI * u2);
            double qc = sqrt(u1) * sin(2 * M_PI * u3), qd = sqrt(u1) * cos(2 * M_PI * u3);
            /* quaternion (w=qd, x=qa, y=qb, z=qc) -> rotation matrix */
            double qw = qd, qx = qa, qy = qb, qz = qc;
            double R[3][3] = {
                {1 - 2 * (qy * qy + qz * qz), 2 * (qx * qy - qz * qw), 2 * (qx * qz + qy * qw)},
                {2 * (qx * qy + qz * qw), 1 - 2 * (qx * qx + qz * qz), 2 * (qy * qz - qx * qw)},
                {2 * (qx * qz - qy * qw), 2 * (qy * qz + qx * qw), 1 - 2 * (qx * qx + qy * qy)}};
            /* monomer geometry, optionally perturbed (learned-guess training set) */
            double r1 = hypot(0.7572, 0.5865), r2 = r1;
            double half = atan2(0.7572, 0.5865); /* half the HOH angle */
            double h1[3] = {0.7572, 0.5865, 0.0}, h2[3] = {-0.7572, 0.5865, 0.0};
            if (oh_jitter > 0 || hoh_jitter_deg > 0) {
                r1 += (2 * xo_uniform(&r) - 1) * oh_jitter;
                r2 += (2 * xo_uniform(&r) - 1) * oh_jitter;
                half += 0.5 * (2 * xo_uniform(&r) - 1) * hoh_jitter_deg * M_PI / 180.0;
                h1[0] = r1 * sin(half); h1[1] = r1 * cos(half); h1[2] = 0;
                h2[0] = -r2 * sin(half); h2[1] = r2 * cos(half); h2[2] = 0;
            }
            double* a = xyz + 9 * w;
            for (int c = 0; c < 3; ++c) a[c] = o[c];
            for (int c = 0; c < 3; ++c) {
                a[3 + c] = o[c] + R[c][0] * h1[0] + R[c][1] * h1[1] + R[c][2] * h1[2];
                a[6 + c] = o[c] + R[c][0] * h2[0] + R[c][1] * h2[1] + R[c][2] * h2[2];
            }
        }
        int ok = 1;
        for (int a = 0; a < n && ok; ++a)
            for (int b = a + 1; b < n; ++b) {
                double dx = xyz[9 * a] - xyz[9 * b], dy = xyz[9 * a + 1] - xyz[9 * b + 1],
                       dz = xyz[9 * a + 2] - xyz[9 * b + 2];
                if (dx * dx + dy * dy + dz * dz < 2.5 * 2.5) { ok = 0; break; }
            }
        if (ok) return attempt;
    }
    return -1;
}

/* Seeded orthonormal C (nbf×nocc, column-major): N(0,1) draws in column-major
 * order, Householder QR, Q with diag(R) >= 0. */
void orc_random_orthonormal(uint64_t seed, size_t nbf, size_t nocc, double* C) {
    xoshiro r;
    xo_seed(&r, seed);
    size_t m = nbf, n = nocc;
    double* A = (double*)malloc(sizeof(double) * m * n);
    double* V = (double*)calloc(m * n, sizeof(double));
    double* rdiag = (double*)malloc(sizeof(double) * n);
    for (size_t i = 0; i < m * n; ++i) A[i] = xo_normal(&r);
    for (size_t k = 0; k < n; ++k) {
        double* a = A + k * m;
        double nrm = 0;
        for (size_t i = k; i < m; ++i) nrm += a[i] * a[i];
        nrm = sqrt(nrm);
        double alpha = a[k] > 0 ? -nrm : nrm;
        double* v = V + k * m;
        for (size_t i = k; i < m; ++i) v[i] = a[i];
        v[k] -= alpha;
        double vn = 0;
        for (size_t i = k; i < m; ++i) vn += v[i] * v[i];
        vn = sqrt(vn);
        if (vn > 0)
            for (size_t i = k; i < m; ++i) v[i] /= vn;
        for (size_t j = k; j < n; ++j) {
            double* aj = A + j * m;
            double d = 0;
            for (size_t i = k; i < m; ++i) d += v[i] * aj[i];
            for (size_t i = k; i < m; ++i) aj[i] -= 2 * d * v[i];
        }
        rdiag[k] = A[k * m + k];
    }
    /* Q = H_0 ... H_{n-1} [I_n; 0] */
    for (size_t j = 0; j < n; ++j) {
        double* q = C + j * m;
        for (size_t i = 0; i < m; ++i) q[i] = (i == j) ? 1.0 : 0.0;
        for (size_t kk = n; kk-- > 0;) {
            double* v = V + kk * m;
            double d = 0;
            for (size_t i = kk; i < m; ++i) d += v[i] * q[i];
            for (size_t i = kk; i < m; ++i) q[i] -= 2 * d * v[i];
        }
        if (rdiag[j] < 0)
            for (size_t i = 0; i < m; ++i) q[i] = -q[i];
    }
    free(A); free(V); free(rdiag);
}

/* ------------------------------------------------------------------------ */
/* a1  vj_upper_with_rimatr_v_sync  (src/scf_io/mod.rs:2490-2561)            */
/*   γ_P = 2 Σ_{upper} B D − Σ_diag B D  (:2516-2529); J_p = s Σ_P B γ_P     */
/* ------------------------------------------------------------------------ */
void orc_vj_upper_rimatr(const double* b, size_t nbf, size_t naux, const double* dm,
                         size_t nspin, double scale, double* vj) {
    const size_t npair = nbf * (nbf + 1) / 2;
    double* gam = (double*)malloc(sizeof(double) * naux);
    double* dup = (double*)malloc(sizeof(double) * npair);
    for (size_t s = 0; s < nspin; ++s) {
        const double* d = dm + s * nbf * nbf;
        for (size_t j = 0; j < nbf; ++j)
            for (size_t i = 0; i <= j; ++i) dup[pidx(i, j)] = d[i + j * nbf]; /* iter_matrixupper */
#pragma omp parallel for schedule(static)
        for (long long P = 0; P < (long long)naux; ++P) {
            const double* m = b + (size_t)P * npair;
            nsum full = {0, 0}, diag = {0, 0};
            for (size_t p = 0; p < npair; ++p) nadd(&full, m[p] * dup[p]);
            for (size_t i = 0; i < nbf; ++i) nadd(&diag, m[pidx(i, i)] * d[i + i * nbf]);
            gam[P] = 2.0 * nval(&full) - nval(&diag);
        }
        double* out = vj + s * npair;
#pragma omp parallel for schedule(static)
        for (long long p = 0; p < (long long)npair; ++p) {
            nsum acc = {0, 0};
            for (size_t P = 0; P < naux; ++P) nadd(&acc, b[P * npair + p] * gam[P]);
            out[p] = nval(&acc) * scale;
        }
    }
    free(gam); free(dup);
}

/* γ only (for tests of the first pass) */
void orc_gamma_rimatr(const double* b, size_t nbf, size_t naux, const double* dm, double* gam) {
    const size_t npair = nbf * (nbf + 1) / 2;
    for (size_t P = 0; P < naux; ++P) {
        const double* m = b + P * npair;
        nsum full = {0, 0}, diag = {0, 0};
        for (size_t j = 0; j < nbf; ++j)
            for (size_t i = 0; i <= j; ++i) nadd(&full, m[pidx(i, j)] * dm[i + j * nbf]);
        for (size_t i = 0; i < nbf; ++i) nadd(&diag, m[pidx(i, i)] * dm[i + i * nbf]);
        gam[P] = 2.0 * nval(&full) - nval(&diag);
    }
}

/* ------------------------------------------------------------------------ */
/* a2  vj_upper_with_ri_v_sync  (src/scf_io/mod.rs:2428-2488)               */
/*   γ_P = Σ_{kl} B_{kl,P} D_{kl} (full, :2453-2457); J upper filter (:2459) */
/* ------------------------------------------------------------------------ */
void orc_vj_upper_ri3fn(const double* b3, size_t nbf, size_t naux, const double* dm,
                        size_t nspin, double scale, double* vj) {
    const size_t npair = nbf * (nbf + 1) / 2, n2 = nbf * nbf;
    double* gam = (double*)malloc(sizeof(double) * naux);
    for (size_t s = 0; s < nspin; ++s) {
        const double* d = dm + s * n2;
#pragma omp parallel for schedule(static)
        for (long long P = 0; P < (long long)naux; ++P) {
            nsum acc = {0, 0};
            const double* m = b3 + (size_t)P * n2;
            for (size_t k = 0; k < n2; ++k) nadd(&acc, m[k] * d[k]);
            gam[P] = nval(&acc);
        }
        double* out = vj + s * npair;
#pragma omp parallel for schedule(static)
        for (long long j = 0; j < (long long)nbf; ++j)
            for (size_t i = 0; i <= (size_t)j; ++i) {
                nsum acc = {0, 0};
                for (size_t P = 0; P < naux; ++P) nadd(&acc, b3[P * n2 + i + (size_t)j * nbf] * gam[P]);
                out[pidx(i, (size_t)j)] = nval(&acc) * scale;
            }
    }
    free(gam);
}

/* ------------------------------------------------------------------------ */
/* K: per-P partial upper triangles summed with compensation over P.        */
/* ------------------------------------------------------------------------ */
typedef void (*kpart_fn)(const void* arg, size_t P, double* kfull /* nbf*nbf */, double* scratch);

static void k_accumulate(size_t nbf, size_t naux, kpart_fn fn, const void* arg, size_t scratch_len,
                         double scale, double* out) {
    const size_t npair = nbf * (nbf + 1) / 2;
    nsum* acc = (nsum*)calloc(npair, sizeof(nsum));
#pragma omp parallel
    {
        nsum* loc = (nsum*)calloc(npair, sizeof(nsum));
        double* kf = (double*)malloc(sizeof(double) * nbf * nbf);
        double* scr = (double*)malloc(sizeof(double) * (scratch_len ? scratch_len : 1));
#pragma omp for schedule(dynamic, 1)
        for (long long P = 0; P < (long long)naux; ++P) {
            fn(arg, (size_t)P, kf, scr);
            for (size_t j = 0; j < nbf; ++j)
                for (size_t i = 0; i <= j; ++i) nadd(&loc[pidx(i, j)], kf[i + j * nbf]); /* upper filter */
        }
#pragma omp critical
        for (size_t p = 0; p < npair; ++p) { nadd(&acc[p], loc[p].s); nadd(&acc[p], loc[p].c); }
        free(loc); free(kf); free(scr);
    }
    for (size_t p = 0; p < npair; ++p) out[p] = nval(&acc[p]) * scale;
    free(acc);
}

/* B_P as a full matrix: packed -> symmetric (upper copied to both triangles,
 * which is what dsymm 'U' reads), or ri3fn column taken as is. */
static void unpack_sym(const double* col, size_t nbf, double* full) {
    for (size_t j = 0; j < nbf; ++j)
        for (size_t i = 0; i <= j; ++i) full[i + j * nbf] = full[j + i * nbf] = col[pidx(i, j)];
}
static void upper_sym(const double* m, size_t nbf, double* full) {
    for (size_t j = 0; j < nbf; ++j)
        for (size_t i = 0; i <= j; ++i) full[i + j * nbf] = full[j + i * nbf] = m[i + j * nbf];
}

/* dense C = A·B (n×k)(k×m), column-major, compensated dot products.  A is
 * transposed into a scratch copy first so that every dot product runs over
 * contiguous memory. */
static void gemm_nn(const double* A, size_t lda, const double* B, size_t ldb, double* C, size_t ldc,
                    size_t n, size_t k, size_t m) {
    double* At = (double*)malloc(sizeof(double) * n * k);
    for (size_t l = 0; l < k; ++l)
        for (size_t i = 0; i < n; ++i) At[l + i * k] = A[i + l * lda];
    for (size_t j = 0; j < m; ++j)
        for (size_t i = 0; i < n; ++i) {
            nsum s = {0, 0};
            const double* a = At + i * k;
            const double* b = B + j * ldb;
            for (size_t l = 0; l < k; ++l) nadd(&s, a[l] * b[l]);
            C[i + j * ldc] = nval(&s);
        }
    free(At);
}

/* a3  vk_upper_with_ri_v_sync (src/scf_io/mod.rs:2758-2834)                 */
/*   M = B_P C[:,0..nw] (:2795); W = M diag(occ) (:2797-2800); K_P = W Mᵀ (:2803) */
typedef struct { const double* b; int packed; size_t nbf, npair, nw; const double* c; const double* occ; } occ_arg;
static void kpart_occ(const void* va, size_t P, double* kf, double* scr) {
    const occ_arg* a = (const occ_arg*)va;
    size_t nbf = a->nbf, nw = a->nw;
    double* bp = scr;              /* nbf*nbf */
    double* M = scr + nbf * nbf;   /* nbf*nw  */
    if (a->packed) unpack_sym(a->b + P * a->npair, nbf, bp);
    else memcpy(bp, a->b + P * nbf * nbf, sizeof(double) * nbf * nbf);
    gemm_nn(bp, nbf, a->c, nbf, M, nbf, nbf, nbf, nw);
    double* Mt = M + nbf * nw;     /* nw*nbf, row o contiguous per μ */
    for (size_t o = 0; o < nw; ++o)
        for (size_t i = 0; i < nbf; ++i) Mt[o + i * nw] = M[i + o * nbf];
    for (size_t j = 0; j < nbf; ++j)
        for (size_t i = 0; i <= j; ++i) {
            nsum s = {0, 0};
            const double* mi = Mt + i * nw;
            const double* mj = Mt + j * nw;
            for (size_t o = 0; o < nw; ++o) nadd(&s, mi[o] * a->occ[o] * mj[o]);
            kf[i + j * nbf] = nval(&s);
        }
}
void orc_vk_upper_occ(const double* b, int packed, size_t nbf, size_t naux, const double* eigv,
                      size_t nstate, const size_t* homo, const double* occupation, size_t nspin,
                      double scale, double* vk) {
    const size_t npair = nbf * (nbf + 1) / 2;
    for (size_t s = 0; s < nspin; ++s) {
        occ_arg a = {b, packed, nbf, npair, homo[s] + 1, eigv + s * nbf * nstate, occupation + s * nstate};
        k_accumulate(nbf, naux, kpart_occ, &a, nbf * nbf + 2 * nbf * (homo[s] + 1), scale, vk + s * npair);
    }
}

/* a4/a5  vk_upper_with_rimatr_use_dm_only_sync (:2703-2756) and
 *        vk_upper_with_ri_v_use_dm_only_sync (:2640-2701):
 *   T = Bsym·D (_dsymm L,U), K_P = T·Bsym (_dsymm R,U), Bsym from the upper
 *   triangle of B_P.  D is used as a general matrix, as dsymm does. */
typedef struct { const double* b; int packed; size_t nbf, npair; const double* d; } dm_arg;
static void kpart_dm(const void* va, size_t P, double* kf, double* scr) {
    const dm_arg* a = (const dm_arg*)va;
    size_t nbf = a->nbf;
    double* bp = scr;
    double* T = scr + nbf * nbf;
    if (a->packed) unpack_sym(a->b + P * a->npair, nbf, bp);
    else upper_sym(a->b + P * nbf * nbf, nbf, bp);
    gemm_nn(bp, nbf, a->d, nbf, T, nbf, nbf, nbf, nbf);
    gemm_nn(T, nbf, bp, nbf, kf, nbf, nbf, nbf, nbf);
}
void orc_vk_upper_dm(const double* b, int packed, size_t nbf, size_t naux, const double* dm,
                     size_t nspin, double scale, double* vk) {
    const size_t npair = nbf * (nbf + 1) / 2;
    for (size_t s = 0; s < nspin; ++s) {
        dm_arg a = {b, packed, nbf, npair, dm + s * nbf * nbf};
        k_accumulate(nbf, naux, kpart_dm, &a, 2 * nbf * nbf, scale, vk + s * npair);
    }
}

/* a12 tail  generate_ri3fn_from_rimatr (src/molecule_io/mod.rs:1725-1744):
 * packed column -> full symmetric nbf×nbf column. */
void orc_ri3fn_from_rimatr(const double* b, size_t nbf, size_t naux, double* b3) {
    const size_t npair = nbf * (nbf + 1) / 2;
    for (size_t P = 0; P < naux; ++P) unpack_sym(b + P * npair, nbf, b3 + P * nbf * nbf);
}

/* a9  generate_density_matrix (src/scf_io/mod.rs:441-478): D = (C·occ)Cᵀ over nw = homo+1 */
void orc_density_matrix(const double* eigv, size_t nbf, size_t nstate, const size_t* homo,
                        const double* occupation, size_t nspin, double* dm) {
    for (size_t s = 0; s < nspin; ++s) {
        const double* c = eigv + s * nbf * nstate;
        const double* occ = occupation + s * nstate;
        size_t nw = homo[s] + 1;
        for (size_t j = 0; j < nbf; ++j)
            for (size_t i = 0; i < nbf; ++i) {
                nsum a = {0, 0};
                for (size_t o = 0; o < nw; ++o) nadd(&a, c[i + o * nbf] * occ[o] * c[j + o * nbf]);
                dm[s * nbf * nbf + i + j * nbf] = nval(&a);
            }
    }
}

/* a7  generate_hf_hamiltonian_ri_v (src/scf_io/mod.rs:1339-1381):
 *   F_s = H + J_0 + J_1 + K_s   (K already carries s_K = −0.5 RHF / −1 UHF, :1347-1350)
 * all packed upper. */
void orc_fock_upper(const double* hcore, const double* vj, const double* vk, size_t npair, size_t nspin,
                    double* fock) {
    for (size_t s = 0; s < nspin; ++s)
        for (size_t p = 0; p < npair; ++p) {
            double f = hcore[p] + vj[p];
            if (nspin == 2) f += vj[npair + p];
            fock[s * npair + p] = f + vk[s * npair + p];
        }
}

/* a8  par_energy_contraction (src/scf_io/mod.rs:1787-1802): for symmetric
 * D (full) and A (packed upper), Σ_{μν} D_{μν} A_{μν} = 2Σ_upper − Σ_diag. */
double orc_energy_contraction(const double* dm, const double* a_upper, size_t nbf) {
    nsum full = {0, 0}, diag = {0, 0};
    for (size_t j = 0; j < nbf; ++j)
        for (size_t i = 0; i <= j; ++i) nadd(&full, dm[i + j * nbf] * a_upper[pidx(i, j)]);
    for (size_t i = 0; i < nbf; ++i) nadd(&diag, dm[i + i * nbf] * a_upper[pidx(i, i)]);
    return 2.0 * nval(&full) - nval(&diag);
}

/* ------------------------------------------------------------------------ */
/* CPU baseline: the reference's structure (task parallelism over P, one     */
/* single-threaded BLAS call per task, serial accumulation of the partials)  */
/* with a host BLAS resolved at run time through its public dgemm symbol.    */
/* ------------------------------------------------------------------------ */
typedef void (*dgemm64_t)(const char*, const char*, const int64_t*, const int64_t*, const int64_t*,
                          const double*, const double*, const int64_t*, const double*, const int64_t*,
                          const double*, double*, const int64_t*, size_t, size_t);
typedef void (*setthreads_t)(int);
static dgemm64_t g_dgemm = 0;
static setthreads_t g_setthreads = 0;

/* Load a BLAS exporting the ILP64 Fortran dgemm under `dgemm_sym` (e.g. the
 * OpenBLAS wheel of numpy: "scipy_dgemm_64_").  Returns 1 on success. */
int orc_blas_load(const char* path, const char* dgemm_sym, const char* setthreads_sym) {
    void* h = dlopen(path, RTLD_NOW | RTLD_LOCAL);
    if (!h) return 0;
    g_dgemm = (dgemm64_t)dlsym(h, dgemm_sym);
    g_setthreads = setthreads_sym ? (setthreads_t)dlsym(h, setthreads_sym) : 0;
    return g_dgemm != 0;
}

static void bl_gemm(char ta, char tb, int64_t m, int64_t n, int64_t k, const double* A, int64_t lda,
                    const double* B, int64_t ldb, double beta, double* C, int64_t ldc) {
    double one = 1.0;
    if (g_dgemm) {
        g_dgemm(&ta, &tb, &m, &n, &k, &one, A, &lda, B, &ldb, &beta, C, &ldc, 1, 1);
        return;
    }
    /* built-in fallback: plain loops (only ta = 'N') */
    for (int64_t j = 0; j < n; ++j)
        for (int64_t i = 0; i < m; ++i) {
            double s = 0;
            for (int64_t l = 0; l < k; ++l) {
                double bv = (tb == 'N') ? B[l + j * ldb] : B[j + l * ldb];
                s += A[i + l * lda] * bv;
            }
            C[i + j * ldc] = beta * C[i + j * ldc] + s;
        }
}

/* J as REST runs it (a1): per-P dot + scaled copy, serial accumulation.
 * `ncol` columns of packed B (a bounded sample of the workload). */
double orc_cpu_baseline_vj(const double* b, size_t nbf, size_t ncol, const double* dm, double* vj,
                           int threads) {
    const size_t npair = nbf * (nbf + 1) / 2;
#ifdef _OPENMP
    double t0 = omp_get_wtime();
    if (threads > 0) omp_set_num_threads(threads);
#endif
    double* dup = (double*)malloc(sizeof(double) * npair);
    for (size_t j = 0; j < nbf; ++j)
        for (size_t i = 0; i <= j; ++i) dup[pidx(i, j)] = dm[i + j * nbf];
    memset(vj, 0, sizeof(double) * npair);
#pragma omp parallel
    {
        double* part = (double*)malloc(sizeof(double) * npair);
#pragma omp for schedule(dynamic, 1)
        for (long long P = 0; P < (long long)ncol; ++P) {
            const double* m = b + (size_t)P * npair;
            double full = 0, diag = 0;
            for (size_t p = 0; p < npair; ++p) full += m[p] * dup[p];
            for (size_t i = 0; i < nbf; ++i) diag += m[pidx(i, i)] * dm[i + i * nbf];
            double g = 2 * full - diag;
            for (size_t p = 0; p < npair; ++p) part[p] = m[p] * g;
#pragma omp critical
            for (size_t p = 0; p < npair; ++p) vj[p] += part[p];
        }
        free(part);
    }
    free(dup);
#ifdef _OPENMP
    return omp_get_wtime() - t0;
#else
    return 0;
#endif
}

/* K as REST's occupied route runs it (a3): per P unpack, dgemm NN, column
 * scaling, dgemm NT, upper filter, serial accumulation. */
double orc_cpu_baseline_vk(const double* b, size_t nbf, size_t ncol, const double* c, size_t nw,
                           const double* occ, double* vk, int threads) {
    const size_t npair = nbf * (nbf + 1) / 2;
#ifdef _OPENMP
    double t0 = omp_get_wtime();
    if (threads > 0) omp_set_num_threads(threads);
#endif
    if (g_setthreads) g_setthreads(1);
    memset(vk, 0, sizeof(double) * npair);
#pragma omp parallel
    {
        double* bp = (double*)malloc(sizeof(double) * nbf * nbf);
        double* M = (double*)malloc(sizeof(double) * nbf * nw);
        double* W = (double*)malloc(sizeof(double) * nbf * nw);
        double* kf = (double*)malloc(sizeof(double) * nbf * nbf);
        double* up = (double*)malloc(sizeof(double) * npair);
#pragma omp for schedule(dynamic, 1)
        for (long long P = 0; P < (long long)ncol; ++P) {
            unpack_sym(b + (size_t)P * npair, nbf, bp);
            bl_gemm('N', 'N', nbf, nw, nbf, bp, nbf, c, nbf, 0.0, M, nbf);
            for (size_t o = 0; o < nw; ++o)
                for (size_t i = 0; i < nbf; ++i) W[i + o * nbf] = M[i + o * nbf] * occ[o];
            bl_gemm('N', 'T', nbf, nbf, nw, W, nbf, M, nbf, 0.0, kf, nbf);
            for (size_t j = 0; j < nbf; ++j)
                for (size_t i = 0; i <= j; ++i) up[pidx(i, j)] = kf[i + j * nbf];
#pragma omp critical
            for (size_t p = 0; p < npair; ++p) vk[p] += up[p];
        }
        free(bp); free(M); free(W); free(kf); free(up);
    }
#ifdef _OPENMP
    return omp_get_wtime() - t0;
#else
    return 0;
#endif
}

/* K as REST's density-only route runs it (a4, vk_upper_with_rimatr_use_dm_only_sync,
 * src/scf_io/mod.rs:2703-2756): per P unpack, T = Bsym·D and K_P = T·Bsym (REST's two
 * dsymm calls, here as dgemm on the unpacked symmetric B_P), upper filter, serial
 * accumulation. */
double orc_cpu_baseline_vk_dm(const double* b, size_t nbf, size_t ncol, const double* dm, double* vk,
                              int threads) {
    const size_t npair = nbf * (nbf + 1) / 2;
#ifdef _OPENMP
    double t0 = omp_get_wtime();
    if (threads > 0) omp_set_num_threads(threads);
#endif
    if (g_setthreads) g_setthreads(1);
    memset(vk, 0, sizeof(double) * npair);
#pragma omp parallel
    {
        double* bp = (double*)malloc(sizeof(double) * nbf * nbf);
        double* T = (double*)malloc(sizeof(double) * nbf * nbf);
        double* kf = (double*)malloc(sizeof(double) * nbf * nbf);
        double* up = (double*)malloc(sizeof(double) * npair);
#pragma omp for schedule(dynamic, 1)
        for (long long P = 0; P < (long long)ncol; ++P) {
            unpack_sym(b + (size_t)P * npair, nbf, bp);
            bl_gemm('N', 'N', nbf, nbf, nbf, bp, nbf, dm, nbf, 0.0, T, nbf);
            bl_gemm('N', 'N', nbf, nbf, nbf, T, nbf, bp, nbf, 0.0, kf, nbf);
            for (size_t j = 0; j < nbf; ++j)
                for (size_t i = 0; i <= j; ++i) up[pidx(i, j)] = kf[i + j * nbf];
#pragma omp critical
            for (size_t p = 0; p < npair; ++p) vk[p] += up[p];
        }
        free(bp); free(T); free(kf); free(up);
    }
#ifdef _OPENMP
    return omp_get_wtime() - t0;
#else
    return 0;
#endif
}

int orc_max_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
