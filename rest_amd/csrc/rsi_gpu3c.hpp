// rsi_gpu3c.hpp — device three-centre integrals (μν|P) for the B producer.
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <vector>

#include "rsi_internal.hpp"

namespace rsi {

// A shell as the device kernels read it; primitives and the Cartesian ->
// spherical rows live in flat arrays at p0 / t0.
struct GShell {
    double A[3];
    int l, nprim, p0, t0, off, pad;
};

constexpr int kGpuMaxLOrb = 2;   // orbital shells up to d
constexpr int kGpuMaxLAux = 4;   // auxiliary shells up to g
constexpr int kGpuClasses = (kGpuMaxLOrb + 1) * (kGpuMaxLOrb + 2) / 2;   // (la >= lb) pairs

struct Gpu3c {
    GShell* orb = nullptr;
    GShell* aux = nullptr;
    double* pexp = nullptr;
    double* pcoef = nullptr;
    double* tsph = nullptr;
    double* boys = nullptr;   // Boys table, rows of Boys::kMTab at T = i·Boys::kStep
    int* auxl = nullptr;      // auxiliary shells grouped by l
    int2* pairs = nullptr;    // screened shell pairs (A.l >= B.l), grouped by slab, then class
    int aux_off[kGpuMaxLAux + 1] = {}, aux_cnt[kGpuMaxLAux + 1] = {};
    std::vector<size_t> cuts;   // slab boundaries in pair columns j
    std::vector<size_t> cls_off, cls_cnt;   // [slab * kGpuClasses + class]
};

bool gpu3c_supported(const Basis& orb, const Basis& aux);
// Upload shells, primitives, transforms and the Boys table; build the pair
// lists of every slab [cuts[k], cuts[k+1]).  Returns an RJK_* code.
int gpu3c_init(Gpu3c* g, const Basis& orb, const Basis& aux, const std::vector<size_t>& cuts, hipStream_t st);
// (μν|P) of slab k into the device array X[naux][nq] (cleared first), on st.
int gpu3c_slab(const Gpu3c& g, size_t k, hipStream_t st, double* X, size_t naux);
void gpu3c_free(Gpu3c* g);

}  // namespace rsi
