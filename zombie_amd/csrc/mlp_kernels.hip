// mlp_kernels.hip -- Fourier-feature MLP control variate on gfx950: forward, backward and
// Adam written by hand, dense layers on MFMA (bf16 inputs, fp32 accumulation).
//
// Network c(x): phi(x) = [sin(B x), cos(B x)] (B: F x d, fixed, N(0, sigma^2) entries)
//               -> L hidden layers of H units with SiLU -> 1 linear output.
// Training minimises 0.5 * mean_b (c(x_b) - y_b)^2 with Adam on fp32 master weights; the
// GEMM operands (weights, activations) are bf16 copies.  The input gradient grad_x c is
// the control variate used by the walk kernel (first step, zero mean).
//
// Dense layers use one tiled kernel: C[M x N] = A[M x K] * B[K x N] with arbitrary element
// strides for A and B, a 64 x 64 block tile (4 waves, each 32 x 32 = 2 x 2 MFMA 16x16x16
// tiles) staged through LDS, and a fused epilogue:
//   FWD       : z = acc + bias[n];  Z = z (fp32);  H = bf16(silu(z))
//   BWD_DATA  : dZprev = bf16(acc * silu'(Zprev))   (gradient w.r.t. the previous pre-activation)
//   BWD_INPUT : dPhi = acc (fp32)                    (gradient w.r.t. the Fourier features)
//   BWD_WEIGHT: atomicAdd(dW, acc)                   (split over the batch: K = rows)
#include <hip/hip_runtime.h>

#include <cfloat>
#include <cmath>

#include "mlp.h"

namespace zg {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint16_t toBf16(float f) {
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)((u >> 16) | ((u & 0xffffu) ? 0x40u : 0u));  // inf/nan
    uint32_t r = u + 0x7fffu + ((u >> 16) & 1u);  // round to nearest even
    return (uint16_t)(r >> 16);
}
__device__ __forceinline__ float fromBf16(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }
__device__ __forceinline__ float silu(float z) { return z / (1.0f + __expf(-z)); }
__device__ __forceinline__ float siluGrad(float z) {
    float s = 1.0f / (1.0f + __expf(-z));
    return s * (1.0f + z * (1.0f - s));
}

enum Epilogue { kFwd = 0, kBwdData = 1, kBwdInput = 2, kBwdWeight = 3 };

struct GemmArgs {
    int M, N, K;
    const uint16_t* A; int64_t sAm, sAk;   // element strides
    const uint16_t* B; int64_t sBk, sBn;
    const float* bias;                     // FWD
    float* Z;                              // FWD: out fp32 [M x N];  BWD_DATA: in, previous pre-activation
    uint16_t* H;                           // FWD: out bf16 [M x N];  BWD_DATA: out dZprev bf16 [M x N]
    float* F;                              // BWD_INPUT: out fp32 [M x N];  BWD_WEIGHT: atomic fp32 [M x N]
    int ldc;                               // row stride of the outputs
};

constexpr int kBM = 64, kBN = 64, kBK = 16;

// MFMA v_mfma_f32_16x16x16_bf16 operand layout (lane l of 64, element j of 4):
//   A[row = l % 16][k = 4 (l / 16) + j],  B[k = 4 (l / 16) + j][col = l % 16],
//   D[row = 4 (l / 16) + j][col = l % 16]
template <int EPI>
__global__ void __launch_bounds__(256) gemmKernel(GemmArgs g) {
    __shared__ uint16_t As[kBM][kBK + 4];
    __shared__ uint16_t Bs[kBK][kBN + 4];
    const int tid = threadIdx.x;
    const int wave = tid >> 6, lane = tid & 63;
    const int m0 = blockIdx.y * kBM, n0 = blockIdx.x * kBN;
    // BWD_WEIGHT splits the reduction (K) over blockIdx.z
    int kBegin = 0, kEnd = g.K;
    if (EPI == kBwdWeight) {
        int chunk = (g.K + gridDim.z - 1) / gridDim.z;
        chunk = (chunk + kBK - 1) / kBK * kBK;
        kBegin = blockIdx.z * chunk;
        kEnd = min(g.K, kBegin + chunk);
    }
    const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
    f32x4 acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++) acc[i][j] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};

    for (int k0 = kBegin; k0 < kEnd; k0 += kBK) {
        // stage A (64 x 16) and B (16 x 64): 1024 elements each, 4 per thread
#pragma unroll
        for (int e = 0; e < 4; e++) {
            int idx = tid + e * 256;
            int r = idx / kBK, c = idx % kBK;
            int gm = m0 + r, gk = k0 + c;
            As[r][c] = (gm < g.M && gk < kEnd) ? g.A[(int64_t)gm * g.sAm + (int64_t)gk * g.sAk] : (uint16_t)0;
            int rb = idx / kBN, cb = idx % kBN;
            int gk2 = k0 + rb, gn = n0 + cb;
            Bs[rb][cb] = (gk2 < kEnd && gn < g.N) ? g.B[(int64_t)gk2 * g.sBk + (int64_t)gn * g.sBn] : (uint16_t)0;
        }
        __syncthreads();
        bf16x4 af[2], bfr[2];
#pragma unroll
        for (int i = 0; i < 2; i++) {
            int r = wm + 16 * i + (lane & 15);
            int kk = 4 * (lane >> 4);
            af[i] = bf16x4{(short)As[r][kk], (short)As[r][kk + 1], (short)As[r][kk + 2], (short)As[r][kk + 3]};
            int c = wn + 16 * i + (lane & 15);
            bfr[i] = bf16x4{(short)Bs[kk][c], (short)Bs[kk + 1][c], (short)Bs[kk + 2][c], (short)Bs[kk + 3][c]};
        }
#pragma unroll
        for (int i = 0; i < 2; i++)
#pragma unroll
            for (int j = 0; j < 2; j++) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(af[i], bfr[j], acc[i][j], 0, 0, 0);
        __syncthreads();
    }
    // epilogue
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++)
#pragma unroll
            for (int e = 0; e < 4; e++) {
                int row = m0 + wm + 16 * i + 4 * (lane >> 4) + e;
                int col = n0 + wn + 16 * j + (lane & 15);
                if (row >= g.M || col >= g.N) continue;
                float v = acc[i][j][e];
                int64_t o = (int64_t)row * g.ldc + col;
                if (EPI == kFwd) {
                    float z = v + g.bias[col];
                    g.Z[o] = z;
                    g.H[o] = toBf16(silu(z));
                } else if (EPI == kBwdData) {
                    g.H[o] = toBf16(v * siluGrad(g.Z[o]));
                } else if (EPI == kBwdInput) {
                    g.F[o] = v;
                } else {
                    atomicAdd(&g.F[o], v);
                }
            }
}

// phi = [sin(B x), cos(B x)] in bf16, [n x 2F]
__global__ void fourierKernel(const float* __restrict__ x, int n, int d, const float* __restrict__ Bf, int F,
                              uint16_t* __restrict__ phi) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int64_t)n * F) return;
    int row = (int)(i / F), f = (int)(i % F);
    float a = 0.0f;
    for (int k = 0; k < d; k++) a += Bf[f * d + k] * x[(int64_t)row * d + k];
    float s, c;
    sincosf(a, &s, &c);
    phi[(int64_t)row * 2 * F + f] = toBf16(s);
    phi[(int64_t)row * 2 * F + F + f] = toBf16(c);
}

// output layer c = H . w + b (fp32 weights); loss gradient dc = (c - y) / n; dZlast =
// bf16(dc * w * silu'(Zlast)); dw, db accumulated with block reductions + atomics
__global__ void __launch_bounds__(256) outputKernel(const uint16_t* __restrict__ Hl, const float* __restrict__ Zl, int n,
                                                    int H, const float* __restrict__ w, const float* __restrict__ b,
                                                    const float* __restrict__ y, float invN, float* __restrict__ out,
                                                    uint16_t* __restrict__ dZl, float* __restrict__ dw,
                                                    float* __restrict__ db, float* __restrict__ loss, int mode) {
    // mode 0: value only; 1: training (needs y); 2: input gradient (dc = 1)
    __shared__ float sdw[256];
    __shared__ float sred[256];
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    float c = 0.0f, dc = 0.0f;
    bool valid = i < n;
    if (valid) {
        float acc = b[0];
        for (int k = 0; k < H; k++) acc += fromBf16(Hl[(int64_t)i * H + k]) * w[k];
        c = acc;
        if (out) out[i] = c;
        if (mode == 1) dc = (c - y[i]) * invN;
        else if (mode == 2) dc = 1.0f;
        if (mode != 0)
            for (int k = 0; k < H; k++)
                dZl[(int64_t)i * H + k] = toBf16(dc * w[k] * siluGrad(Zl[(int64_t)i * H + k]));
    }
    if (mode != 1) return;
    // loss and bias gradient
    sred[threadIdx.x] = valid ? 0.5f * (c - y[i]) * (c - y[i]) * invN : 0.0f;
    sdw[threadIdx.x] = dc;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) { sred[threadIdx.x] += sred[threadIdx.x + s]; sdw[threadIdx.x] += sdw[threadIdx.x + s]; }
        __syncthreads();
    }
    if (threadIdx.x == 0) { if (loss) atomicAdd(loss, sred[0]); atomicAdd(db, sdw[0]); }
    // dw[k] = sum_i dc_i H[i][k]
    for (int k = 0; k < H; k++) {
        __syncthreads();
        sred[threadIdx.x] = valid ? dc * fromBf16(Hl[(int64_t)i * H + k]) : 0.0f;
        __syncthreads();
        for (int s = 128; s > 0; s >>= 1) {
            if (threadIdx.x < s) sred[threadIdx.x] += sred[threadIdx.x + s];
            __syncthreads();
        }
        if (threadIdx.x == 0) atomicAdd(&dw[k], sred[0]);
    }
}

// db[n] = sum_rows dZ[row][n] (bf16 dZ), split over blocks; columns are strided over the
// block when N exceeds it
__global__ void __launch_bounds__(256) biasGradKernel(const uint16_t* __restrict__ dZ, int rows, int N,
                                                      float* __restrict__ db) {
    int lanesPerCol = max(1, (int)blockDim.x / N);
    int sub = threadIdx.x / N;
    if (sub >= lanesPerCol) return;
    for (int col = threadIdx.x % N; col < N; col += (int)blockDim.x) {
        float acc = 0.0f;
        for (int r = blockIdx.x * lanesPerCol + sub; r < rows; r += gridDim.x * lanesPerCol)
            acc += fromBf16(dZ[(int64_t)r * N + col]);
        atomicAdd(&db[col], acc);
        if (N <= (int)blockDim.x) break;
    }
}

// grad_x c = sum_f dPhi_sin[f] cos(B_f x) B_f - dPhi_cos[f] sin(B_f x) B_f
__global__ void inputGradKernel(const float* __restrict__ x, int n, int d, const float* __restrict__ Bf, int F,
                                const float* __restrict__ dPhi, float* __restrict__ grad) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float gx[3] = {0.0f, 0.0f, 0.0f};
    for (int f = 0; f < F; f++) {
        float a = 0.0f;
        for (int k = 0; k < d; k++) a += Bf[f * d + k] * x[(int64_t)i * d + k];
        float s, c;
        sincosf(a, &s, &c);
        float coef = dPhi[(int64_t)i * 2 * F + f] * c - dPhi[(int64_t)i * 2 * F + F + f] * s;
        for (int k = 0; k < d; k++) gx[k] += coef * Bf[f * d + k];
    }
    for (int k = 0; k < d; k++) grad[(int64_t)i * d + k] = gx[k];
}

// Adam (Kingma & Ba) on fp32 master weights; refreshes the bf16 copies of the weight matrices
__global__ void adamKernel(float* __restrict__ theta, float* __restrict__ m, float* __restrict__ v,
                           const float* __restrict__ grad, int64_t n, float lr, float b1, float b2, float eps,
                           float c1, float c2, const int64_t* __restrict__ bfIndex, uint16_t* __restrict__ bfCopy) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float g = grad[i];
    float mi = b1 * m[i] + (1.0f - b1) * g;
    float vi = b2 * v[i] + (1.0f - b2) * g * g;
    m[i] = mi; v[i] = vi;
    float t = theta[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    theta[i] = t;
    int64_t bi = bfIndex[i];
    if (bi >= 0) bfCopy[bi] = toBf16(t);
}

__global__ void castKernel(const float* __restrict__ theta, int64_t n, const int64_t* __restrict__ bfIndex,
                           uint16_t* __restrict__ bfCopy) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int64_t bi = bfIndex[i];
    if (bi >= 0) bfCopy[bi] = toBf16(theta[i]);
}

__global__ void zeroKernel(float* __restrict__ p, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.0f;
}

// ------------------------------------------------------------------------------------
// host side
// ------------------------------------------------------------------------------------
static unsigned blocks1d(int64_t n, int bs = 256) { return (unsigned)((n + bs - 1) / bs); }

static hipError_t gemm(int epi, const GemmArgs& g, int splitK, hipStream_t s) {
    dim3 grid((g.N + kBN - 1) / kBN, (g.M + kBM - 1) / kBM, epi == kBwdWeight ? splitK : 1);
    switch (epi) {
    case kFwd: hipLaunchKernelGGL(gemmKernel<kFwd>, grid, dim3(256), 0, s, g); break;
    case kBwdData: hipLaunchKernelGGL(gemmKernel<kBwdData>, grid, dim3(256), 0, s, g); break;
    case kBwdInput: hipLaunchKernelGGL(gemmKernel<kBwdInput>, grid, dim3(256), 0, s, g); break;
    default: hipLaunchKernelGGL(gemmKernel<kBwdWeight>, grid, dim3(256), 0, s, g); break;
    }
    return hipGetLastError();
}

hipError_t mlpZero(float* p, int64_t n, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(zeroKernel, dim3(blocks1d(n)), dim3(256), 0, s, p, n);
    return hipGetLastError();
}

hipError_t mlpCast(MlpDevice& m, hipStream_t s) {
    hipLaunchKernelGGL(castKernel, dim3(blocks1d(m.nParams)), dim3(256), 0, s, m.theta, m.nParams, m.bfIndex, m.wBf);
    return hipGetLastError();
}

// forward over rows [0, n) of x; fills phi, Z_l, H_l and (optionally) the output values
hipError_t mlpForward(MlpDevice& m, const float* x, int n, float* out, hipStream_t s) {
    const MlpShape& sh = m.shape;
    int F = sh.nFreq, twoF = 2 * F, Hd = sh.hidden;
    hipLaunchKernelGGL(fourierKernel, dim3(blocks1d((int64_t)n * F)), dim3(256), 0, s, x, n, sh.inDim, m.freq, F, m.phi);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    for (int l = 0; l < sh.nHidden; l++) {
        GemmArgs g{};
        int K = l == 0 ? twoF : Hd;
        g.M = n; g.N = Hd; g.K = K;
        g.A = l == 0 ? m.phi : m.act[l - 1]; g.sAm = K; g.sAk = 1;
        g.B = m.wBf + m.wOffset[l]; g.sBk = 1; g.sBn = K;          // B[k][n] = W[n][k]
        g.bias = m.theta + m.bOffset[l];
        g.Z = m.pre[l]; g.H = m.act[l]; g.ldc = Hd;
        e = gemm(kFwd, g, 1, s);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

// backward from the output (mode 1: loss vs y; mode 2: dc/dc = 1 for the input gradient)
hipError_t mlpBackward(MlpDevice& m, const float* y, int n, int mode, float* loss, float* out, hipStream_t s) {
    const MlpShape& sh = m.shape;
    int F = sh.nFreq, twoF = 2 * F, Hd = sh.hidden, L = sh.nHidden;
    const float* wOut = m.theta + m.wOffset[L];
    const float* bOut = m.theta + m.bOffset[L];
    float* dwOut = m.grad + m.wOffset[L];
    float* dbOut = m.grad + m.bOffset[L];
    hipLaunchKernelGGL(outputKernel, dim3(blocks1d(n)), dim3(256), 0, s, m.act[L - 1], m.pre[L - 1], n, Hd, wOut, bOut,
                       y, 1.0f / (float)n, out, m.dz[0], dwOut, dbOut, loss, mode);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    int cur = 0;
    for (int l = L - 1; l >= 0; l--) {
        int K = l == 0 ? twoF : Hd;
        const uint16_t* prevAct = l == 0 ? m.phi : m.act[l - 1];
        if (mode == 1) {
            // dW_l = dZ_l^T . H_{l-1}   ([Hd x n] . [n x K]), split over the batch
            GemmArgs g{};
            g.M = Hd; g.N = K; g.K = n;
            g.A = m.dz[cur]; g.sAm = 1; g.sAk = Hd;
            g.B = prevAct; g.sBk = K; g.sBn = 1;
            g.F = m.grad + m.wOffset[l]; g.ldc = K;
            int split = std::max(1, std::min(256, n / 512));
            e = gemm(kBwdWeight, g, split, s);
            if (e != hipSuccess) return e;
            hipLaunchKernelGGL(biasGradKernel, dim3(std::max(1, std::min(512, n / 64))), dim3(256), 0, s, m.dz[cur], n,
                               Hd, m.grad + m.bOffset[l]);
            e = hipGetLastError();
            if (e != hipSuccess) return e;
        }
        // dH_{l-1} = dZ_l . W_l ([n x Hd] . [Hd x K]); through SiLU' for hidden layers
        GemmArgs g{};
        g.M = n; g.N = K; g.K = Hd;
        g.A = m.dz[cur]; g.sAm = Hd; g.sAk = 1;
        g.B = m.wBf + m.wOffset[l]; g.sBk = K; g.sBn = 1;          // B[k][n] = W[k][n] (W is Hd x K)
        g.ldc = K;
        if (l > 0) {
            g.Z = m.pre[l - 1]; g.H = m.dz[cur ^ 1];
            e = gemm(kBwdData, g, 1, s);
            cur ^= 1;
        } else if (mode == 2) {
            g.F = m.dphi;
            e = gemm(kBwdInput, g, 1, s);
        } else {
            e = hipSuccess;
        }
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

hipError_t mlpInputGrad(MlpDevice& m, const float* x, int n, float* grad, hipStream_t s) {
    hipLaunchKernelGGL(inputGradKernel, dim3(blocks1d(n)), dim3(256), 0, s, x, n, m.shape.inDim, m.freq, m.shape.nFreq,
                       m.dphi, grad);
    return hipGetLastError();
}

hipError_t mlpAdam(MlpDevice& m, float lr, hipStream_t s) {
    m.step++;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    float c1 = 1.0f - powf(b1, (float)m.step), c2 = 1.0f - powf(b2, (float)m.step);
    hipLaunchKernelGGL(adamKernel, dim3(blocks1d(m.nParams)), dim3(256), 0, s, m.theta, m.adamM, m.adamV, m.grad,
                       m.nParams, lr, b1, b2, eps, c1, c2, m.bfIndex, m.wBf);
    return hipGetLastError();
}

}  // namespace zg
