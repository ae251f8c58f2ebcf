// rsi_internal.hpp — shells and bases of the B producer (shared by the host
// integral drivers and the device B build).
#pragma once
#include <string>
#include <vector>

namespace rsi {

constexpr double kPairCut = 1e-16;   // primitive pairs with |c_a c_b| exp(-μ R²) below this are skipped

struct Shell {
    int l = 0, nprim = 0, atom = 0, nsph = 0, off = 0;
    double A[3] = {0, 0, 0};
    std::vector<double> a;   // exponents
    std::vector<double> c;   // contraction coefficients × radial normalisation of each primitive
    std::vector<double> T;   // (2l+1) × ncart Cartesian -> spherical rows, scaled to unit-norm functions
};

struct Basis {
    std::vector<Shell> sh;
    size_t nbf = 0;
};

void set_error(const std::string& m);
const std::string& last_error();
bool finish_shell(Shell& s);
void pair_1e(const Shell& A, const Shell& B, size_t ncharge, const double* charge, const double* cxyz, double* S,
             double* T, double* V);
void aux_2c(const Shell& C, const Shell& D, double* out);
void pair_3c(const Shell& A, const Shell& B, const Shell& C, double* out, std::vector<double>& scratch);
double pair_bound(const Shell& A, const Shell& B);
void int3c_slab(const Basis& orb, const Basis& aux, size_t j0, size_t j1, double* out);

inline bool on_shell_boundary(const Basis& b, size_t j) {
    if (j == 0 || j == b.nbf) return true;
    for (const Shell& s : b.sh)
        if ((size_t)s.off == j) return true;
    return false;
}

}  // namespace rsi

struct rsi_basis : rsi::Basis {};
