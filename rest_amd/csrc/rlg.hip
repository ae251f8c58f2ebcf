// rlg.hip — the pair-block MLP of the learned initial guess: forward pass,
// masked-MSE loss, backward pass and Adam, all on the device (include/rest_lg.h).
//
// Activations, gradients and the master weights are FP32.  Every GEMM stages
// its operands in LDS as BF16 (round to nearest even) and multiplies them on
// the BF16 matrix cores (v_mfma_f32_16x16x32_bf16) with FP32 accumulation.
// Layout: row-major activations [batch][features], weights W_l [in][out]
// (Y = X W + b), SiLU after every hidden layer, linear output.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/rest_lg.h"

struct rlg_net {
    int device = 0;
    hipStream_t own = nullptr, stream = nullptr;
    std::vector<int> dims;   // n_in, hidden..., n_out
    size_t max_batch = 0, nparam = 0;
    float* params = nullptr;   // W_0, b_0, W_1, b_1, ... (FP32 master copy)
    float* grads = nullptr;    // same layout
    float* adam_m = nullptr;
    float* adam_v = nullptr;
    std::vector<size_t> w_off, b_off;
    std::vector<float*> act;   // act[0] = input copy pointer (not owned), act[l+1] = layer l output
    std::vector<float*> pre;   // pre-activation of every layer
    float* g0 = nullptr;       // backward activation gradients (ping-pong)
    float* g1 = nullptr;
    float* split = nullptr;    // split-K partials of the weight gradients
    size_t split_cap = 0;
    double* part = nullptr;    // loss partial sums
    long long step = 0;
    std::string err;
};

namespace {

typedef __bf16 bf8 __attribute__((ext_vector_type(8)));
typedef float f4 __attribute__((ext_vector_type(4)));

constexpr int kTM = 64, kTN = 64, kTK = 32, kPad = 8;
constexpr int kLossBlocks = 512;

int fail(rlg_net* n, int code, const std::string& m) {
    if (n) n->err = m;
    return code;
}

#define RLG_HIP(net, call)                                                                              \
    do {                                                                                                \
        hipError_t e_ = (call);                                                                         \
        if (e_ != hipSuccess) return fail((net), RLG_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

// C_z[m][n] = Σ_{k in split z} A(m,k) B(k,n),  A(m,k) = A[m*sam + k*sak], B(k,n) = B[k*sbk + n*sbn];
// split z covers k in [z*kchunk, min(K, (z+1)*kchunk)) and writes C + z*M*ldc.
__global__ __launch_bounds__(256) void k_gemm(const float* __restrict__ A, long sam, long sak,
                                              const float* __restrict__ B, long sbk, long sbn,
                                              float* __restrict__ C, long ldc, int M, int N, int K, int kchunk) {
    __shared__ __bf16 As[kTM][kTK + kPad];
    __shared__ __bf16 Bt[kTN][kTK + kPad];   // B tile stored transposed: Bt[n][k]
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int m0 = blockIdx.y * kTM, n0 = blockIdx.x * kTN;
    const int kbeg = blockIdx.z * kchunk, kend = min(K, kbeg + kchunk);
    C += (size_t)blockIdx.z * M * ldc;
    const int wm = w >> 1, wn = w & 1;
    f4 acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = f4{0.f, 0.f, 0.f, 0.f};
    for (int k0 = kbeg; k0 < kend; k0 += kTK) {
#pragma unroll
        for (int i = 0; i < 8; ++i) {   // A tile 64 x 32
            const int idx = tid + 256 * i, r = idx >> 5, kk = idx & 31;
            const int gm = m0 + r, gk = k0 + kk;
            const float v = (gm < M && gk < kend) ? A[(size_t)gm * sam + (size_t)gk * sak] : 0.f;
            As[r][kk] = (__bf16)v;
        }
#pragma unroll
        for (int i = 0; i < 8; ++i) {   // B tile 32 x 64
            const int idx = tid + 256 * i, c = idx & 63, kk = idx >> 6;
            const int gn = n0 + c, gk = k0 + kk;
            const float v = (gn < N && gk < kend) ? B[(size_t)gk * sbk + (size_t)gn * sbn] : 0.f;
            Bt[c][kk] = (__bf16)v;
        }
        __syncthreads();
        bf8 a[2], b[2];
#pragma unroll
        for (int i = 0; i < 2; ++i) {
            const int r = wm * 32 + i * 16 + (lane & 15);
#pragma unroll
            for (int e = 0; e < 8; ++e) a[i][e] = As[r][8 * (lane >> 4) + e];
        }
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            const int c = wn * 32 + j * 16 + (lane & 15);
#pragma unroll
            for (int e = 0; e < 8; ++e) b[j][e] = Bt[c][8 * (lane >> 4) + e];
        }
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
        __syncthreads();
    }
    // D[4(l/16)+r][l%16] of each 16x16 block
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int gm = m0 + wm * 32 + i * 16 + 4 * (lane >> 4) + r;
                const int gn = n0 + wn * 32 + j * 16 + (lane & 15);
                if (gm < M && gn < N) C[(size_t)gm * ldc + gn] = acc[i][j][r];
            }
}

__device__ inline float sigmoid(float z) { return 1.f / (1.f + __expf(-z)); }

// Z += b (row broadcast); A = SiLU(Z) or Z
__global__ void k_bias_act(float* __restrict__ Z, const float* __restrict__ b, size_t M, int N, int silu,
                           float* __restrict__ A) {
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= M * N) return;
    const float z = Z[t] + b[t % N];
    Z[t] = z;
    A[t] = silu ? z * sigmoid(z) : z;
}

// dZ = G ⊙ SiLU'(Z)
__global__ void k_silu_bwd(const float* __restrict__ G, const float* __restrict__ Z, size_t n, float* __restrict__ dZ) {
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const float z = Z[t], s = sigmoid(z);
    dZ[t] = G[t] * s * (1.f + z * (1.f - s));
}

// partial[2b] = Σ mask·(y − t)², partial[2b+1] = Σ mask   (FP64 sums)
__global__ void k_loss_sums(const float* __restrict__ Y, const float* __restrict__ T, const float* __restrict__ Mk,
                            size_t n, double* __restrict__ partial) {
    __shared__ double s0[256], s1[256];
    double a = 0.0, c = 0.0;
    for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (size_t)gridDim.x * blockDim.x) {
        const double d = (double)Y[t] - (double)T[t];
        a += (double)Mk[t] * d * d;
        c += (double)Mk[t];
    }
    s0[threadIdx.x] = a;
    s1[threadIdx.x] = c;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) {
            s0[threadIdx.x] += s0[threadIdx.x + off];
            s1[threadIdx.x] += s1[threadIdx.x + off];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        partial[2 * blockIdx.x] = s0[0];
        partial[2 * blockIdx.x + 1] = s1[0];
    }
}

// G = 2 mask (y − t) / Σmask
__global__ void k_loss_grad(const float* __restrict__ Y, const float* __restrict__ T, const float* __restrict__ Mk,
                            size_t n, float scale, float* __restrict__ G) {
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) G[t] = scale * Mk[t] * (Y[t] - T[t]);
}

// gb[n] = Σ_m dZ[m][n]
__global__ void k_colsum(const float* __restrict__ dZ, size_t M, int N, float* __restrict__ gb) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= N) return;
    float s = 0.f;
    for (size_t m = 0; m < M; ++m) s += dZ[m * N + c];
    gb[c] = s;
}

// out[i] = Σ_z part[z*count + i]
__global__ void k_sum_splits(const float* __restrict__ part, int splits, size_t count, float* __restrict__ out) {
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= count) return;
    float s = 0.f;
    for (int z = 0; z < splits; ++z) s += part[(size_t)z * count + t];
    out[t] = s;
}

__global__ void k_adam(float* __restrict__ w, const float* __restrict__ g, float* __restrict__ m,
                       float* __restrict__ v, size_t n, float lr, float b1, float b2, float eps, float bc1,
                       float bc2) {
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const float gi = g[t];
    const float mi = b1 * m[t] + (1.f - b1) * gi;
    const float vi = b2 * v[t] + (1.f - b2) * gi * gi;
    m[t] = mi;
    v[t] = vi;
    w[t] -= lr * (mi / bc1) / (sqrtf(vi / bc2) + eps);
}

unsigned blocks(size_t n, unsigned b = 256) { return (unsigned)((n + b - 1) / b); }

hipError_t gemm(hipStream_t st, const float* A, long sam, long sak, const float* B, long sbk, long sbn, float* C,
                long ldc, int M, int N, int K, int splits) {
    const int kchunk = ((K + splits - 1) / splits + kTK - 1) / kTK * kTK;
    dim3 grid(blocks(N, kTN), blocks(M, kTM), splits);
    hipLaunchKernelGGL(k_gemm, grid, dim3(256), 0, st, A, sam, sak, B, sbk, sbn, C, ldc, M, N, K, kchunk);
    return hipGetLastError();
}

int nlayers(const rlg_net* n) { return (int)n->dims.size() - 1; }

// forward over `batch` rows of x; output in n->act.back()
int forward(rlg_net* n, const float* x, size_t batch) {
    const int L = nlayers(n);
    n->act[0] = const_cast<float*>(x);
    for (int l = 0; l < L; ++l) {
        const int din = n->dims[l], dout = n->dims[l + 1];
        RLG_HIP(n, gemm(n->stream, n->act[l], din, 1, n->params + n->w_off[l], dout, 1, n->pre[l], dout, (int)batch,
                        dout, din, 1));
        hipLaunchKernelGGL(k_bias_act, dim3(blocks(batch * dout)), dim3(256), 0, n->stream, n->pre[l],
                           n->params + n->b_off[l], batch, dout, l + 1 < L ? 1 : 0, n->act[l + 1]);
        RLG_HIP(n, hipGetLastError());
    }
    return RLG_OK;
}

int check_batch(rlg_net* n, size_t batch) {
    if (batch == 0 || batch > n->max_batch)
        return fail(n, RLG_ERR_ARG, "batch must be 1.." + std::to_string(n->max_batch));
    return RLG_OK;
}

}  // namespace

extern "C" {

int rlg_create(int device, size_t n_in, size_t n_hidden, size_t n_hidden_layers, size_t n_out, size_t max_batch,
               rlg_net** out) {
    if (!out || n_in == 0 || n_hidden == 0 || n_out == 0 || max_batch == 0 || n_hidden_layers > 16 ||
        n_in > 65536 || n_hidden > 65536 || n_out > 65536 || max_batch > (size_t(1) << 24))
        return RLG_ERR_ARG;
    *out = nullptr;
    rlg_net* n = new rlg_net();
    n->device = device;
    n->dims.push_back((int)n_in);
    for (size_t i = 0; i < n_hidden_layers; ++i) n->dims.push_back((int)n_hidden);
    n->dims.push_back((int)n_out);
    n->max_batch = max_batch;
    const int L = nlayers(n);
    size_t off = 0, widest = 0;
    for (int l = 0; l < L; ++l) {
        n->w_off.push_back(off);
        off += (size_t)n->dims[l] * n->dims[l + 1];
        n->b_off.push_back(off);
        off += n->dims[l + 1];
        widest = std::max(widest, (size_t)std::max(n->dims[l], n->dims[l + 1]));
    }
    n->nparam = off;
    auto alloc = [&](float** p, size_t cnt) { return hipMalloc((void**)p, cnt * sizeof(float)) == hipSuccess; };
    bool ok = hipSetDevice(device) == hipSuccess && hipStreamCreate(&n->own) == hipSuccess;
    n->stream = n->own;
    ok = ok && alloc(&n->params, off) && alloc(&n->grads, off) && alloc(&n->adam_m, off) && alloc(&n->adam_v, off);
    n->act.assign(L + 1, nullptr);
    n->pre.assign(L, nullptr);
    for (int l = 0; l < L && ok; ++l) {
        ok = alloc(&n->pre[l], max_batch * n->dims[l + 1]) && alloc(&n->act[l + 1], max_batch * n->dims[l + 1]);
    }
    ok = ok && alloc(&n->g0, max_batch * widest) && alloc(&n->g1, max_batch * widest);
    ok = ok && hipMalloc((void**)&n->part, 2 * kLossBlocks * sizeof(double)) == hipSuccess;
    // split-K partials for the largest weight gradient
    size_t wmax = 0;
    for (int l = 0; l < L; ++l) wmax = std::max(wmax, (size_t)n->dims[l] * n->dims[l + 1]);
    n->split_cap = 16 * wmax;
    ok = ok && alloc(&n->split, n->split_cap);
    ok = ok && hipMemsetAsync(n->params, 0, off * sizeof(float), n->stream) == hipSuccess &&
         hipMemsetAsync(n->adam_m, 0, off * sizeof(float), n->stream) == hipSuccess &&
         hipMemsetAsync(n->adam_v, 0, off * sizeof(float), n->stream) == hipSuccess &&
         hipStreamSynchronize(n->stream) == hipSuccess;
    if (!ok) {
        rlg_destroy(n);
        return RLG_ERR_NOMEM;
    }
    *out = n;
    return RLG_OK;
}

int rlg_destroy(rlg_net* n) {
    if (!n) return RLG_OK;
    (void)hipSetDevice(n->device);
    if (n->stream) (void)hipStreamSynchronize(n->stream);
    for (float* p : {n->params, n->grads, n->adam_m, n->adam_v, n->g0, n->g1, n->split})
        if (p) (void)hipFree(p);
    for (float* p : n->pre)
        if (p) (void)hipFree(p);
    for (size_t l = 1; l < n->act.size(); ++l)
        if (n->act[l]) (void)hipFree(n->act[l]);
    if (n->part) (void)hipFree(n->part);
    if (n->own) (void)hipStreamDestroy(n->own);
    delete n;
    return RLG_OK;
}

const char* rlg_last_error(const rlg_net* n) { return n ? n->err.c_str() : "NULL network"; }

size_t rlg_num_params(const rlg_net* n) { return n ? n->nparam : 0; }

int rlg_set_stream(rlg_net* n, void* stream) {
    if (!n) return RLG_ERR_ARG;
    n->stream = stream ? (hipStream_t)stream : n->own;
    return RLG_OK;
}

int rlg_set_params(rlg_net* n, const float* host) {
    if (!n || !host) return RLG_ERR_ARG;
    RLG_HIP(n, hipSetDevice(n->device));
    RLG_HIP(n, hipMemcpyAsync(n->params, host, n->nparam * sizeof(float), hipMemcpyHostToDevice, n->stream));
    RLG_HIP(n, hipStreamSynchronize(n->stream));
    return RLG_OK;
}

int rlg_get_params(rlg_net* n, float* host) {
    if (!n || !host) return RLG_ERR_ARG;
    RLG_HIP(n, hipSetDevice(n->device));
    RLG_HIP(n, hipMemcpyAsync(host, n->params, n->nparam * sizeof(float), hipMemcpyDeviceToHost, n->stream));
    RLG_HIP(n, hipStreamSynchronize(n->stream));
    return RLG_OK;
}

int rlg_forward(rlg_net* n, const float* x_dev, size_t batch, float* y_dev) {
    if (!n || !x_dev || !y_dev) return fail(n, RLG_ERR_ARG, "rlg_forward: NULL argument");
    int rc = check_batch(n, batch);
    if (rc) return rc;
    RLG_HIP(n, hipSetDevice(n->device));
    rc = forward(n, x_dev, batch);
    if (rc) return rc;
    RLG_HIP(n, hipMemcpyAsync(y_dev, n->act.back(), batch * n->dims.back() * sizeof(float), hipMemcpyDeviceToDevice,
                              n->stream));
    RLG_HIP(n, hipStreamSynchronize(n->stream));
    return RLG_OK;
}

int rlg_loss_grad(rlg_net* n, const float* x_dev, const float* t_dev, const float* mask_dev, size_t batch,
                  double* loss, float* grad_host) {
    if (!n || !x_dev || !t_dev || !mask_dev) return fail(n, RLG_ERR_ARG, "rlg_loss_grad: NULL argument");
    int rc = check_batch(n, batch);
    if (rc) return rc;
    RLG_HIP(n, hipSetDevice(n->device));
    rc = forward(n, x_dev, batch);
    if (rc) return rc;
    const int L = nlayers(n), nout = n->dims.back();
    const size_t ny = batch * nout;
    hipLaunchKernelGGL(k_loss_sums, dim3(kLossBlocks), dim3(256), 0, n->stream, n->act.back(), t_dev, mask_dev, ny,
                       n->part);
    RLG_HIP(n, hipGetLastError());
    std::vector<double> part(2 * kLossBlocks);
    RLG_HIP(n, hipMemcpyAsync(part.data(), n->part, part.size() * sizeof(double), hipMemcpyDeviceToHost, n->stream));
    RLG_HIP(n, hipStreamSynchronize(n->stream));
    double se = 0.0, msum = 0.0;
    for (int b = 0; b < kLossBlocks; ++b) {
        se += part[2 * b];
        msum += part[2 * b + 1];
    }
    if (!(msum > 0.0)) return fail(n, RLG_ERR_ARG, "rlg_loss_grad: the mask is empty");
    if (loss) *loss = se / msum;
    float* G = n->g0;
    float* Gn = n->g1;
    hipLaunchKernelGGL(k_loss_grad, dim3(blocks(ny)), dim3(256), 0, n->stream, n->act.back(), t_dev, mask_dev, ny,
                       (float)(2.0 / msum), G);
    RLG_HIP(n, hipGetLastError());
    for (int l = L - 1; l >= 0; --l) {
        const int din = n->dims[l], dout = n->dims[l + 1];
        float* dZ = G;
        if (l + 1 < L) {   // through the SiLU of this hidden layer
            hipLaunchKernelGGL(k_silu_bwd, dim3(blocks(batch * dout)), dim3(256), 0, n->stream, G, n->pre[l],
                               batch * dout, Gn);
            RLG_HIP(n, hipGetLastError());
            dZ = Gn;
            std::swap(G, Gn);   // G now holds dZ; Gn is free
        }
        // gW = A_lᵀ dZ  (din × dout, reduction over the batch, split over rows of the batch)
        const size_t wcount = (size_t)din * dout;
        int splits = (int)std::min<size_t>(16, std::max<size_t>(1, batch / 512));
        splits = (int)std::min<size_t>(splits, n->split_cap / wcount);
        RLG_HIP(n, gemm(n->stream, n->act[l], 1, din, dZ, dout, 1, n->split, dout, din, dout, (int)batch, splits));
        hipLaunchKernelGGL(k_sum_splits, dim3(blocks(wcount)), dim3(256), 0, n->stream, n->split, splits, wcount,
                           n->grads + n->w_off[l]);
        RLG_HIP(n, hipGetLastError());
        hipLaunchKernelGGL(k_colsum, dim3(blocks(dout)), dim3(256), 0, n->stream, dZ, batch, dout,
                           n->grads + n->b_off[l]);
        RLG_HIP(n, hipGetLastError());
        if (l > 0) {   // G_{l} = dZ W_lᵀ  (batch × din)
            RLG_HIP(n, gemm(n->stream, dZ, dout, 1, n->params + n->w_off[l], 1, dout, Gn, din, (int)batch, din, dout,
                            1));
            std::swap(G, Gn);
        }
    }
    if (grad_host)
        RLG_HIP(n, hipMemcpyAsync(grad_host, n->grads, n->nparam * sizeof(float), hipMemcpyDeviceToHost, n->stream));
    RLG_HIP(n, hipStreamSynchronize(n->stream));
    return RLG_OK;
}

int rlg_adam_step(rlg_net* n, double lr, double beta1, double beta2, double eps) {
    if (!n) return RLG_ERR_ARG;
    if (!(lr > 0) || !(beta1 >= 0 && beta1 < 1) || !(beta2 >= 0 && beta2 < 1) || !(eps > 0))
        return fail(n, RLG_ERR_ARG, "rlg_adam_step: need lr > 0, 0 <= beta < 1, eps > 0");
    RLG_HIP(n, hipSetDevice(n->device));
    ++n->step;
    const float bc1 = (float)(1.0 - std::pow(beta1, (double)n->step));
    const float bc2 = (float)(1.0 - std::pow(beta2, (double)n->step));
    hipLaunchKernelGGL(k_adam, dim3(blocks(n->nparam)), dim3(256), 0, n->stream, n->params, n->grads, n->adam_m,
                       n->adam_v, n->nparam, (float)lr, (float)beta1, (float)beta2, (float)eps, bc1, bc2);
    RLG_HIP(n, hipGetLastError());
    RLG_HIP(n, hipStreamSynchronize(n->stream));
    return RLG_OK;
}

int rlg_reset_optimizer(rlg_net* n) {
    if (!n) return RLG_ERR_ARG;
    RLG_HIP(n, hipSetDevice(n->device));
    RLG_HIP(n, hipMemsetAsync(n->adam_m, 0, n->nparam * sizeof(float), n->stream));
    RLG_HIP(n, hipMemsetAsync(n->adam_v, 0, n->nparam * sizeof(float), n->stream));
    RLG_HIP(n, hipStreamSynchronize(n->stream));
    n->step = 0;
    return RLG_OK;
}

int rlg_gemm_bf16(rlg_net* n, const float* A, long sam, long sak, const float* B, long sbk, long sbn, float* C,
                  long ldc, int M, int N, int K) {
    if (!n || !A || !B || !C || M <= 0 || N <= 0 || K <= 0 || ldc < N)
        return fail(n, RLG_ERR_ARG, "rlg_gemm_bf16: bad argument");
    RLG_HIP(n, hipSetDevice(n->device));
    RLG_HIP(n, gemm(n->stream, A, sam, sak, B, sbk, sbn, C, ldc, M, N, K, 1));
    RLG_HIP(n, hipStreamSynchronize(n->stream));
    return RLG_OK;
}

}  // extern "C"
