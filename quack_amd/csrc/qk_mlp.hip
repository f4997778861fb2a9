// Per-orbital MLP that predicts evGW quasiparticle shifts (warm start of evGW).
//
// Network: nin -> width (SiLU) -> ... (depth hidden layers) -> 1, FP32 weights.
// Parameters live in one flat FP32 buffer, layer by layer: W_l (out x in,
// row-major) then b_l (out).  Training: FP32 forward/backward kernels, MSE
// loss, Adam, all on the device.  Inference: either the FP32 kernels or a
// fused BF16 kernel on matrix cores (v_mfma_f32_16x16x32_bf16, FP32
// accumulation) with the activations of 16 orbitals per wavefront kept in LDS.
// No counterpart in QuAcK.
#include <cmath>
#include <cstring>
#include <new>
#include <random>
#include <vector>

#include "qk_common.h"

namespace qk {
namespace {

thread_local std::string g_mlp_error;

__device__ inline float sigm(float z) { return 1.f / (1.f + __expf(-z)); }

// Z[b][o] = sum_i X[b][i] W[o][i] + bias[o]; A = silu(Z) if act
__global__ void k_lin_fwd(const float* __restrict__ X, const float* __restrict__ W, const float* __restrict__ bias,
                          float* __restrict__ Z, float* __restrict__ A, int B, int nin, int nout, int act) {
    long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (t >= (long long)B * nout) return;
    int b = (int)(t / nout), o = (int)(t % nout);
    const float* x = X + (long long)b * nin;
    const float* w = W + (long long)o * nin;
    float s = bias[o];
    for (int i = 0; i < nin; ++i) s += x[i] * w[i];
    Z[t] = s;
    if (A) A[t] = act ? s * sigm(s) : s;
}

// loss partials and dY = 2 (Y - T) / B
__global__ void k_mse(const float* __restrict__ Y, const float* __restrict__ T, float* __restrict__ dY, int B,
                      float* __restrict__ loss) {
    __shared__ float sh[256];
    float acc = 0.f;
    for (int b = threadIdx.x; b < B; b += blockDim.x) {
        float d = Y[b] - T[b];
        acc += d * d;
        dY[b] = 2.f * d / B;
    }
    sh[threadIdx.x] = acc;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) loss[0] = sh[0] / B;
}

// dZ = dA * silu'(Z), silu'(z) = s(z) (1 + z (1 - s(z)))
__global__ void k_silu_bwd(const float* __restrict__ Z, float* __restrict__ dA, long long n) {
    long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (t < n) {
        float z = Z[t], s = sigm(z);
        dA[t] *= s * (1.f + z * (1.f - s));
    }
}

// dW[o][i] = sum_b dZ[b][o] X[b][i] ; db[o] = sum_b dZ[b][o] (i == nin slot)
__global__ void k_lin_bwd_w(const float* __restrict__ dZ, const float* __restrict__ X, float* __restrict__ dW,
                            float* __restrict__ db, int B, int nin, int nout) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nout * (nin + 1)) return;
    int o = t / (nin + 1), i = t % (nin + 1);
    float s = 0.f;
    if (i < nin) {
        for (int b = 0; b < B; ++b) s += dZ[(long long)b * nout + o] * X[(long long)b * nin + i];
        dW[(long long)o * nin + i] = s;
    } else {
        for (int b = 0; b < B; ++b) s += dZ[(long long)b * nout + o];
        db[o] = s;
    }
}

// dX[b][i] = sum_o dZ[b][o] W[o][i]
__global__ void k_lin_bwd_x(const float* __restrict__ dZ, const float* __restrict__ W, float* __restrict__ dX, int B,
                            int nin, int nout) {
    long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (t >= (long long)B * nin) return;
    int b = (int)(t / nin), i = (int)(t % nin);
    float s = 0.f;
    for (int o = 0; o < nout; ++o) s += dZ[(long long)b * nout + o] * W[(long long)o * nin + i];
    dX[t] = s;
}

__global__ void k_adam(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m, float* __restrict__ v,
                       int n, float lr, float b1, float b2, float c1, float c2, float eps) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) {
        float gi = g[t];
        float mi = b1 * m[t] + (1.f - b1) * gi;
        float vi = b2 * v[t] + (1.f - b2) * gi * gi;
        m[t] = mi;
        v[t] = vi;
        p[t] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    }
}

// rows of X gathered by an index list (mini-batches)
__global__ void k_gather(const float* __restrict__ X, const float* __restrict__ T, const int* __restrict__ idx,
                         float* __restrict__ Xb, float* __restrict__ Tb, int B, int nin) {
    long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (t < (long long)B * nin) {
        int b = (int)(t / nin), i = (int)(t % nin);
        Xb[t] = X[(long long)idx[b] * nin + i];
    }
    if (t < B) Tb[t] = T[idx[t]];
}

// ---- fused BF16 inference ------------------------------------------------
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f4 __attribute__((ext_vector_type(4)));
constexpr int MAXW = 256;

// Packed weights: hidden layer l as bf16 [out][kpad_l] (kpad multiple of 32),
// biases FP32; last layer FP32 (w_last[width], b_last).  One wavefront per 16
// orbitals; activations of the 16 rows in LDS as bf16 [16][MAXW].
__global__ __launch_bounds__(256) void k_mlp_infer_bf16(const float* __restrict__ X, int n, int nin, int kin,
                                                         int width, int depth, const __bf16* __restrict__ Wp,
                                                         const float* __restrict__ bp, const float* __restrict__ wl,
                                                         float bl, float* __restrict__ Y) {
    __shared__ __bf16 act[4][16][MAXW];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    // rows past n compute on zeros and are not stored (no early exit: the
    // block synchronises below)
    const int row0 = (blockIdx.x * 4 + wave) * 16;
    __bf16(*h)[MAXW] = act[wave];
    // input rows, zero-padded to kin
    for (int t = lane; t < 16 * kin; t += 64) {
        int r = t / kin, k = t % kin;
        float x = (row0 + r < n && k < nin) ? X[(long long)(row0 + r) * nin + k] : 0.f;
        h[r][k] = (__bf16)x;
    }
    __syncthreads();
    const __bf16* W = Wp;
    const float* bias = bp;
    int kcur = kin;
    const int l16 = lane & 15, lk = (lane >> 4) * 8;
    for (int l = 0; l < depth; ++l) {
        f4 out[MAXW / 16];
        const int ntile = width / 16;
        for (int jt = 0; jt < ntile; ++jt) {
            f4 c = {0.f, 0.f, 0.f, 0.f};
            const __bf16* wrow = W + (long long)(jt * 16 + l16) * kcur + lk;
            for (int k = 0; k < kcur; k += 32) {
                bf16x8 a, b;
                for (int u = 0; u < 8; ++u) a[u] = h[l16][k + lk + u];
                b = *reinterpret_cast<const bf16x8*>(wrow + k);
                c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
            }
            out[jt] = c;
        }
        __syncthreads();
        // D(i, j) in lane j + 16 (i / 4), register i % 4: row i, neuron jt*16 + j
        for (int jt = 0; jt < ntile; ++jt)
            for (int r = 0; r < 4; ++r) {
                int i = 4 * (lane >> 4) + r, j = jt * 16 + l16;
                float z = out[jt][r] + bias[j];
                h[i][j] = (__bf16)(z * sigm(z));
            }
        __syncthreads();
        W += (long long)width * kcur;
        bias += width;
        kcur = width;
    }
    // last layer in FP32: lane handles row lane/4, quarter of the width
    const int r = lane >> 2, q = lane & 3;
    float s = 0.f;
    for (int j = q; j < width; j += 4) s += (float)h[r][j] * wl[j];
    s += __shfl_xor(s, 1, 64);
    s += __shfl_xor(s, 2, 64);
    if (q == 0 && row0 + r < n) Y[row0 + r] = s + bl;
}

inline int blocks(long long n, int bs = 256) { return (int)((n + bs - 1) / bs); }

template <class T>
T* dmalloc(size_t n) {
    void* p = nullptr;
    if (hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(T)) != hipSuccess) {
        (void)hipGetLastError();
        throw Error{QK_E_ALLOC, "hipMalloc failed in the MLP"};
    }
    return static_cast<T*>(p);
}

}  // namespace
}  // namespace qk

struct qk_mlp_ctx {
    int nin = 0, width = 0, depth = 0, np = 0;
    std::vector<int> lin, lout;        // layer sizes
    std::vector<long long> woff, boff;  // offsets in the flat buffer
    float* p = nullptr;                 // parameters (device)
    float *g = nullptr, *m = nullptr, *v = nullptr;
    hipStream_t s = nullptr;
    std::string err;
    long long adam_t = 0;
};

namespace qk {
namespace {

using Mlp = qk_mlp_ctx;

template <class F>
int guard(Mlp* h, F&& fn) {
    try {
        fn();
        return QK_OK;
    } catch (const Error& e) {
        (h ? h->err : g_mlp_error) = e.msg;
        return e.code;
    } catch (...) {
        (h ? h->err : g_mlp_error) = "unknown error";
        return QK_E_HIP;
    }
}

struct Acts {
    std::vector<float*> Z, A;  // per layer: pre-activation, activation (B x out)
    float* Y = nullptr;
    int B = 0;
    void alloc(const Mlp* h, int B_) {
        B = B_;
        for (size_t l = 0; l < h->lout.size(); ++l) {
            Z.push_back(dmalloc<float>((size_t)B * h->lout[l]));
            A.push_back(dmalloc<float>((size_t)B * h->lout[l]));
        }
    }
    ~Acts() {
        for (auto p : Z) (void)hipFree(p);
        for (auto p : A) (void)hipFree(p);
    }
};

void forward(Mlp* h, const float* X, Acts& a) {
    const float* in = X;
    int L = (int)h->lout.size();
    for (int l = 0; l < L; ++l) {
        int act = l < L - 1;
        long long n = (long long)a.B * h->lout[l];
        hipLaunchKernelGGL(k_lin_fwd, dim3(blocks(n)), dim3(256), 0, h->s, in, h->p + h->woff[l], h->p + h->boff[l],
                           a.Z[l], a.A[l], a.B, h->lin[l], h->lout[l], act);
        in = a.A[l];
    }
    QK_HIP(hipGetLastError());
}

// loss and gradient (into h->g) of the MSE for targets T
void backward(Mlp* h, const float* X, const float* T, Acts& a, float* dloss) {
    int L = (int)h->lout.size();
    float* dY = dmalloc<float>(a.B);
    std::vector<float*> dA(L, nullptr);
    try {
        hipLaunchKernelGGL(k_mse, dim3(1), dim3(256), 0, h->s, a.A[L - 1], T, dY, a.B, dloss);
        float* dZ = dY;
        for (int l = L - 1; l >= 0; --l) {
            const float* in = l > 0 ? a.A[l - 1] : X;
            int nin = h->lin[l], nout = h->lout[l];
            hipLaunchKernelGGL(k_lin_bwd_w, dim3(blocks((long long)nout * (nin + 1))), dim3(256), 0, h->s, dZ, in,
                               h->g + h->woff[l], h->g + h->boff[l], a.B, nin, nout);
            if (l > 0) {
                dA[l - 1] = dmalloc<float>((size_t)a.B * nin);
                hipLaunchKernelGGL(k_lin_bwd_x, dim3(blocks((long long)a.B * nin)), dim3(256), 0, h->s, dZ,
                                   h->p + h->woff[l], dA[l - 1], a.B, nin, nout);
                hipLaunchKernelGGL(k_silu_bwd, dim3(blocks((long long)a.B * nin)), dim3(256), 0, h->s, a.Z[l - 1],
                                   dA[l - 1], (long long)a.B * nin);
                dZ = dA[l - 1];
            }
        }
        QK_HIP(hipGetLastError());
        QK_HIP(hipStreamSynchronize(h->s));
    } catch (...) {
        (void)hipStreamSynchronize(h->s);
        (void)hipFree(dY);
        for (auto p : dA) if (p) (void)hipFree(p);
        throw;
    }
    (void)hipFree(dY);
    for (auto p : dA) if (p) (void)hipFree(p);
}

}  // namespace
}  // namespace qk

using namespace qk;

extern "C" {

const char* qk_mlp_last_error(qk_mlp_handle h) { return h ? h->err.c_str() : g_mlp_error.c_str(); }

int qk_mlp_create(qk_mlp_handle* out, int nin, int width, int depth, unsigned long long seed) {
    if (!out || nin <= 0 || width <= 0 || width > MAXW || width % 16 || depth < 1 || nin > MAXW) return QK_E_ARG;
    *out = nullptr;
    Mlp* h = new (std::nothrow) Mlp();
    if (!h) return QK_E_ALLOC;
    int rc = guard(nullptr, [&] {
        h->nin = nin; h->width = width; h->depth = depth;
        int prev = nin;
        long long off = 0;
        for (int l = 0; l <= depth; ++l) {
            int o = l < depth ? width : 1;
            h->lin.push_back(prev);
            h->lout.push_back(o);
            h->woff.push_back(off);
            off += (long long)o * prev;
            h->boff.push_back(off);
            off += o;
            prev = o;
        }
        h->np = (int)off;
        QK_HIP(hipStreamCreateWithFlags(&h->s, hipStreamNonBlocking));
        h->p = dmalloc<float>(h->np);
        h->g = dmalloc<float>(h->np);
        h->m = dmalloc<float>(h->np);
        h->v = dmalloc<float>(h->np);
        // He-uniform initialisation (fixed seed), zero biases
        std::vector<float> init(h->np, 0.f);
        std::mt19937_64 rng(seed);
        for (int l = 0; l <= depth; ++l) {
            float lim = std::sqrt(6.f / h->lin[l]);
            std::uniform_real_distribution<float> u(-lim, lim);
            for (long long k = 0; k < (long long)h->lout[l] * h->lin[l]; ++k) init[h->woff[l] + k] = u(rng);
        }
        QK_HIP(hipMemcpy(h->p, init.data(), sizeof(float) * h->np, hipMemcpyHostToDevice));
        QK_HIP(hipMemset(h->m, 0, sizeof(float) * h->np));
        QK_HIP(hipMemset(h->v, 0, sizeof(float) * h->np));
    });
    if (rc != QK_OK) {
        if (h->s) (void)hipStreamDestroy(h->s);
        for (float* p : {h->p, h->g, h->m, h->v}) if (p) (void)hipFree(p);
        delete h;
        return rc;
    }
    *out = h;
    return QK_OK;
}

int qk_mlp_destroy(qk_mlp_handle h) {
    if (!h) return QK_E_ARG;
    (void)hipStreamSynchronize(h->s);
    for (float* p : {h->p, h->g, h->m, h->v}) if (p) (void)hipFree(p);
    (void)hipStreamDestroy(h->s);
    delete h;
    return QK_OK;
}

int qk_mlp_nparams(qk_mlp_handle h, int* n) {
    if (!h || !n) return QK_E_ARG;
    *n = h->np;
    return QK_OK;
}

int qk_mlp_get_params(qk_mlp_handle h, float* out) {
    if (!h || !out) return QK_E_ARG;
    return guard(h, [&] { QK_HIP(hipMemcpy(out, h->p, sizeof(float) * h->np, hipMemcpyDeviceToHost)); });
}

int qk_mlp_set_params(qk_mlp_handle h, const float* in) {
    if (!h || !in) return QK_E_ARG;
    return guard(h, [&] {
        QK_HIP(hipMemcpy(h->p, in, sizeof(float) * h->np, hipMemcpyHostToDevice));
        QK_HIP(hipMemset(h->m, 0, sizeof(float) * h->np));
        QK_HIP(hipMemset(h->v, 0, sizeof(float) * h->np));
        h->adam_t = 0;
    });
}

int qk_mlp_forward(qk_mlp_handle h, const float* X, int n, float* Y, int bf16) {
    if (!h || !X || !Y || n < 0) return QK_E_ARG;
    if (n == 0) return QK_OK;
    return guard(h, [&] {
        float* dX = dmalloc<float>((size_t)n * h->nin);
        float* dY = nullptr;
        std::vector<void*> tmp{dX};
        try {
            QK_HIP(hipMemcpyAsync(dX, X, sizeof(float) * n * h->nin, hipMemcpyHostToDevice, h->s));
            if (!bf16) {
                Acts a;
                a.alloc(h, n);
                forward(h, dX, a);
                QK_HIP(hipMemcpyAsync(Y, a.A.back(), sizeof(float) * n, hipMemcpyDeviceToHost, h->s));
                QK_HIP(hipStreamSynchronize(h->s));
            } else {
                // pack hidden weights as bf16 [out][kpad]
                int kin = (h->nin + 31) / 32 * 32;
                std::vector<float> hp(h->np);
                QK_HIP(hipMemcpy(hp.data(), h->p, sizeof(float) * h->np, hipMemcpyDeviceToHost));
                std::vector<__bf16> wp;
                std::vector<float> bp;
                for (int l = 0; l < h->depth; ++l) {
                    int kp = l == 0 ? kin : h->width;
                    for (int o = 0; o < h->width; ++o)
                        for (int k = 0; k < kp; ++k)
                            wp.push_back((__bf16)(k < h->lin[l] ? hp[h->woff[l] + (long long)o * h->lin[l] + k] : 0.f));
                    for (int o = 0; o < h->width; ++o) bp.push_back(hp[h->boff[l] + o]);
                }
                __bf16* dW = dmalloc<__bf16>(wp.size());
                float* db = dmalloc<float>(bp.size());
                float* dwl = dmalloc<float>(h->width);
                dY = dmalloc<float>(n);
                tmp.insert(tmp.end(), {dW, db, dwl, dY});
                QK_HIP(hipMemcpyAsync(dW, wp.data(), sizeof(__bf16) * wp.size(), hipMemcpyHostToDevice, h->s));
                QK_HIP(hipMemcpyAsync(db, bp.data(), sizeof(float) * bp.size(), hipMemcpyHostToDevice, h->s));
                QK_HIP(hipMemcpyAsync(dwl, hp.data() + h->woff[h->depth], sizeof(float) * h->width,
                                      hipMemcpyHostToDevice, h->s));
                float bl = hp[h->boff[h->depth]];
                int nb = (n + 63) / 64;
                hipLaunchKernelGGL(k_mlp_infer_bf16, dim3(nb), dim3(256), 0, h->s, dX, n, h->nin, kin, h->width,
                                   h->depth, dW, db, dwl, bl, dY);
                QK_HIP(hipGetLastError());
                QK_HIP(hipMemcpyAsync(Y, dY, sizeof(float) * n, hipMemcpyDeviceToHost, h->s));
                QK_HIP(hipStreamSynchronize(h->s));
            }
        } catch (...) {
            (void)hipStreamSynchronize(h->s);
            for (void* p : tmp) (void)hipFree(p);
            throw;
        }
        for (void* p : tmp) (void)hipFree(p);
    });
}

int qk_mlp_loss_grad(qk_mlp_handle h, const float* X, const float* T, int n, float* loss, float* grad) {
    if (!h || !X || !T || n <= 0) return QK_E_ARG;
    return guard(h, [&] {
        float* dX = dmalloc<float>((size_t)n * h->nin);
        float* dT = dmalloc<float>(n);
        float* dl = dmalloc<float>(1);
        try {
            QK_HIP(hipMemcpyAsync(dX, X, sizeof(float) * n * h->nin, hipMemcpyHostToDevice, h->s));
            QK_HIP(hipMemcpyAsync(dT, T, sizeof(float) * n, hipMemcpyHostToDevice, h->s));
            Acts a;
            a.alloc(h, n);
            forward(h, dX, a);
            backward(h, dX, dT, a, dl);
            if (loss) QK_HIP(hipMemcpy(loss, dl, sizeof(float), hipMemcpyDeviceToHost));
            if (grad) QK_HIP(hipMemcpy(grad, h->g, sizeof(float) * h->np, hipMemcpyDeviceToHost));
        } catch (...) {
            (void)hipFree(dX); (void)hipFree(dT); (void)hipFree(dl);
            throw;
        }
        (void)hipFree(dX); (void)hipFree(dT); (void)hipFree(dl);
    });
}

// Mini-batch Adam over (X, T) for `epochs` passes; loss_per_epoch (may be
// NULL) receives the mean batch loss of each epoch.
int qk_mlp_train(qk_mlp_handle h, const float* X, const float* T, int n, int epochs, int batch, float lr,
                 unsigned long long seed, float* loss_per_epoch) {
    if (!h || !X || !T || n <= 0 || epochs < 0 || batch <= 0) return QK_E_ARG;
    return guard(h, [&] {
        int B = std::min(batch, n);
        float *dX = dmalloc<float>((size_t)n * h->nin), *dT = dmalloc<float>(n);
        float *Xb = dmalloc<float>((size_t)B * h->nin), *Tb = dmalloc<float>(B), *dl = dmalloc<float>(1);
        int* didx = dmalloc<int>(B);
        std::vector<void*> tmp{dX, dT, Xb, Tb, dl, didx};
        try {
            QK_HIP(hipMemcpyAsync(dX, X, sizeof(float) * n * h->nin, hipMemcpyHostToDevice, h->s));
            QK_HIP(hipMemcpyAsync(dT, T, sizeof(float) * n, hipMemcpyHostToDevice, h->s));
            Acts a;
            a.alloc(h, B);
            std::mt19937_64 rng(seed);
            std::vector<int> perm(n);
            for (int i = 0; i < n; ++i) perm[i] = i;
            const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
            for (int ep = 0; ep < epochs; ++ep) {
                std::shuffle(perm.begin(), perm.end(), rng);
                double esum = 0.0;
                int nb = 0;
                for (int s0 = 0; s0 + B <= n; s0 += B) {
                    QK_HIP(hipMemcpyAsync(didx, perm.data() + s0, sizeof(int) * B, hipMemcpyHostToDevice, h->s));
                    hipLaunchKernelGGL(k_gather, dim3(blocks((long long)B * h->nin)), dim3(256), 0, h->s, dX, dT,
                                       didx, Xb, Tb, B, h->nin);
                    forward(h, Xb, a);
                    backward(h, Xb, Tb, a, dl);
                    ++h->adam_t;
                    float c1 = 1.f - std::pow(b1, (float)h->adam_t), c2 = 1.f - std::pow(b2, (float)h->adam_t);
                    hipLaunchKernelGGL(k_adam, dim3(blocks(h->np)), dim3(256), 0, h->s, h->p, h->g, h->m, h->v, h->np,
                                       lr, b1, b2, c1, c2, eps);
                    float lb = 0.f;
                    QK_HIP(hipMemcpyAsync(&lb, dl, sizeof(float), hipMemcpyDeviceToHost, h->s));
                    QK_HIP(hipStreamSynchronize(h->s));
                    esum += lb;
                    ++nb;
                }
                if (loss_per_epoch) loss_per_epoch[ep] = (float)(esum / std::max(nb, 1));
            }
        } catch (...) {
            (void)hipStreamSynchronize(h->s);
            for (void* p : tmp) (void)hipFree(p);
            throw;
        }
        for (void* p : tmp) (void)hipFree(p);
    });
}

}  // extern "C"
