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
// strides for A and B, a 64 x 64 block tile (4 waves, each 32 x 32 = 2 x 2 MFMA 16x16x32
// tiles) staged through LDS 32 deep in k, and a fused epilogue:
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
    float* dbias;                          // BWD_WEIGHT (nullable): atomic fp32 [M], row sums of A
    int ldc;                               // row stride of the outputs
};

constexpr int kBM = 64, kBN = 64, kBK = 32;
constexpr int kLdsK = kBK + 8;   // padded LDS row (bf16): 80 B, a multiple of 16 B for 128-bit reads

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

// 8 elements of a strided bf16 operand starting at element `base`, `step` elements apart: one
// 128-bit load when they are contiguous, aligned and inside the valid range, else element by element
__device__ __forceinline__ uint4 load8(const uint16_t* __restrict__ p, int64_t base, int64_t step, int valid) {
    if (step == 1 && valid >= 8 && ((reinterpret_cast<uintptr_t>(p + base) & 15) == 0))
        return *reinterpret_cast<const uint4*>(p + base);
    uint16_t v[8];
#pragma unroll
    for (int j = 0; j < 8; j++) v[j] = j < valid ? p[base + (int64_t)j * step] : (uint16_t)0;
    return make_uint4((uint32_t)v[0] | ((uint32_t)v[1] << 16), (uint32_t)v[2] | ((uint32_t)v[3] << 16),
                      (uint32_t)v[4] | ((uint32_t)v[5] << 16), (uint32_t)v[6] | ((uint32_t)v[7] << 16));
}

// stores 8 loaded elements into an LDS tile kept [row][k]: along k when the load ran along k,
// else down 8 rows at one k
__device__ __forceinline__ void store8(uint16_t (*T)[kLdsK], int row, int k, bool alongK, uint4 v) {
    if (alongK) {
        *reinterpret_cast<uint4*>(&T[row][k]) = v;
    } else {
        const uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
        for (int j = 0; j < 8; j++) T[row + j][k] = (uint16_t)(w[j >> 1] >> (16 * (j & 1)));
    }
}

// MFMA v_mfma_f32_16x16x32_bf16 (gfx950) operand layout, lane l of 64, element j of 8:
//   A[row = l % 16][k = 8 (l / 16) + j],  B[k = 8 (l / 16) + j][col = l % 16],
//   D[row = 4 (l / 16) + j][col = l % 16] (j of 4)
// Both operands sit in LDS with k contiguous (A as [m][k], B transposed as [n][k]), so each
// lane fetches its 8 elements with one 128-bit LDS read.  Global operands are read 8 elements at a
// time along whichever of their two dimensions is contiguous (128-bit loads when aligned).
template <int EPI>
__global__ void __launch_bounds__(256) gemmKernel(GemmArgs g) {
    __shared__ __attribute__((aligned(16))) uint16_t As[kBM][kLdsK];
    __shared__ __attribute__((aligned(16))) uint16_t Bs[kBN][kLdsK];
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
    // how each operand tile is read: along k (A row-major / B column-major) or along m / n
    const bool aAlongK = g.sAk == 1;
    const bool bAlongK = g.sBk == 1;
    const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
    f32x4 acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++) acc[i][j] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    // BWD_WEIGHT with dbias: the blocks of the first column tile also sum the rows of their A
    // tiles over k (thread t: row t / 4, k quarter t % 4 of each LDS tile)
    const bool rowSums = EPI == kBwdWeight && g.dbias && blockIdx.x == 0;
    float rsum = 0.0f;

    for (int k0 = kBegin; k0 < kEnd; k0 += kBK) {
        // stage A (64 x 32) and B (32 x 64): 2048 elements each, 8 per thread
        {
            int r, kk, valid;
            int64_t base, step;
            if (aAlongK) {
                r = tid >> 2; kk = (tid & 3) * 8;
                int gm = m0 + r, gk = k0 + kk;
                valid = gm < g.M ? max(0, min(8, kEnd - gk)) : 0;
                base = (int64_t)gm * g.sAm + (int64_t)gk; step = 1;
            } else {
                kk = tid >> 3; r = (tid & 7) * 8;
                int gm = m0 + r, gk = k0 + kk;
                valid = gk < kEnd ? max(0, min(8, g.M - gm)) : 0;
                base = (int64_t)gm * g.sAm + (int64_t)gk * g.sAk; step = g.sAm;
            }
            store8(As, r, kk, aAlongK, load8(g.A, base, step, valid));
        }
        {
            int c, kk, valid;
            int64_t base, step;
            if (bAlongK) {
                c = tid >> 2; kk = (tid & 3) * 8;
                int gn = n0 + c, gk = k0 + kk;
                valid = gn < g.N ? max(0, min(8, kEnd - gk)) : 0;
                base = (int64_t)gk + (int64_t)gn * g.sBn; step = 1;
            } else {
                kk = tid >> 3; c = (tid & 7) * 8;
                int gn = n0 + c, gk = k0 + kk;
                valid = gk < kEnd ? max(0, min(8, g.N - gn)) : 0;
                base = (int64_t)gk * g.sBk + (int64_t)gn * g.sBn; step = g.sBn;
            }
            store8(Bs, c, kk, bAlongK, load8(g.B, base, step, valid));
        }
        __syncthreads();
        if (rowSums) {
            const uint4 v = *reinterpret_cast<const uint4*>(&As[tid >> 2][8 * (tid & 3)]);
            const uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
            for (int j = 0; j < 4; j++) rsum += fromBf16((uint16_t)(w[j] & 0xffffu)) + fromBf16((uint16_t)(w[j] >> 16));
        }
        bf16x8 af[2], bfr[2];
        const int kk = 8 * (lane >> 4);
#pragma unroll
        for (int i = 0; i < 2; i++) {
            af[i] = __builtin_bit_cast(bf16x8, *reinterpret_cast<const uint4*>(&As[wm + 16 * i + (lane & 15)][kk]));
            bfr[i] = __builtin_bit_cast(bf16x8, *reinterpret_cast<const uint4*>(&Bs[wn + 16 * i + (lane & 15)][kk]));
        }
#pragma unroll
        for (int i = 0; i < 2; i++)
#pragma unroll
            for (int j = 0; j < 2; j++) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[i], bfr[j], acc[i][j], 0, 0, 0);
        __syncthreads();
    }
    if (rowSums) {
        rsum += __shfl_xor(rsum, 1);
        rsum += __shfl_xor(rsum, 2);
        if ((tid & 3) == 0 && m0 + (tid >> 2) < g.M) atomicAdd(&g.dbias[m0 + (tid >> 2)], rsum);
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

// Output layer and loss, one wave per row: lane l of a wave owns the columns k = l + 64 j
// (j < NCH, hidden <= 64 NCH), so a row is read as contiguous 128-byte pieces, and each lane
// accumulates the gradient sums of its own columns over the rows its wave visits; the four waves
// of a block are then combined in LDS and added to the results with one atomic per column and
// block.  A wave keeps kRowsInFlight rows' loads in flight.
constexpr int kRowBlocks = 512;   // grid cap of the row-parallel kernels
constexpr int kRowsInFlight = 4;

// adds the per-lane column partials of the block's four waves to dst[0, H)
template <int NCH>
__device__ __forceinline__ void blockColumnAtomics(const float (&part)[NCH], int H, float* __restrict__ dst) {
    __shared__ float sc[4][NCH * 64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int j = 0; j < NCH; j++) sc[wave][lane + 64 * j] = part[j];
    __syncthreads();
    for (int k = threadIdx.x; k < H; k += blockDim.x) atomicAdd(&dst[k], sc[0][k] + sc[1][k] + sc[2][k] + sc[3][k]);
    __syncthreads();
}

// c = H . w + b (fp32 weights); mode 0: values only; mode 1 (training, needs y): loss
// sum 0.5 (c - y)^2 / n, dc = (c - y) / n, dZlast = bf16(dc * w * silu'(Zlast)), and the
// gradients dw = sum_i dc_i H_i, db = sum_i dc_i; mode 2 (input gradient): dZlast with dc = 1.
// Launch: 256 threads, any grid.
template <int NCH>
__global__ void __launch_bounds__(256) outputKernel(const uint16_t* __restrict__ Hl, const float* __restrict__ Zl, int n,
                                                    int H, const float* __restrict__ w, const float* __restrict__ b,
                                                    const float* __restrict__ y, float invN, float* __restrict__ out,
                                                    uint16_t* __restrict__ dZl, float* __restrict__ dw,
                                                    float* __restrict__ db,
                                                    float* __restrict__ loss, int mode) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    float wk[NCH], dwp[NCH];
#pragma unroll
    for (int j = 0; j < NCH; j++) {
        const int k = lane + 64 * j;
        wk[j] = k < H ? w[k] : 0.0f;
        dwp[j] = 0.0f;
    }
    const float b0 = b[0];
    float lossp = 0.0f, dcp = 0.0f;
    const int stride = gridDim.x * 4;
    for (int r0 = blockIdx.x * 4 + wave; r0 < n; r0 += kRowsInFlight * stride) {
        float h[kRowsInFlight][NCH], z[kRowsInFlight][NCH], acc[kRowsInFlight], yv[kRowsInFlight];
#pragma unroll
        for (int u = 0; u < kRowsInFlight; u++) {
            const int i = r0 + u * stride;
            const bool ok = i < n;
            const uint16_t* hr = Hl + (int64_t)i * H;
            const float* zr = Zl + (int64_t)i * H;
            acc[u] = 0.0f;
#pragma unroll
            for (int j = 0; j < NCH; j++) {
                const int k = lane + 64 * j;
                h[u][j] = ok && k < H ? fromBf16(hr[k]) : 0.0f;
                z[u][j] = ok && k < H && mode != 0 ? zr[k] : 0.0f;
                acc[u] += h[u][j] * wk[j];
            }
            yv[u] = ok && mode == 1 ? y[i] : 0.0f;
        }
#pragma unroll
        for (int u = 0; u < kRowsInFlight; u++) {
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) acc[u] += __shfl_xor(acc[u], off);
        }
#pragma unroll
        for (int u = 0; u < kRowsInFlight; u++) {
            const int i = r0 + u * stride;
            if (i >= n) break;
            const float c = acc[u] + b0;
            if (out && lane == 0) out[i] = c;
            if (mode == 0) continue;
            float dc = 1.0f;
            if (mode == 1) {
                const float e = c - yv[u];
                dc = e * invN;
                lossp += 0.5f * e * e * invN;
                dcp += dc;
            }
            uint16_t* dr = dZl + (int64_t)i * H;
#pragma unroll
            for (int j = 0; j < NCH; j++) {
                const int k = lane + 64 * j;
                if (k < H) dr[k] = toBf16(dc * wk[j] * siluGrad(z[u][j]));
                dwp[j] += dc * h[u][j];
            }
        }
    }
    if (mode != 1) return;
    __shared__ float sl[4][2];
    if (lane == 0) { sl[wave][0] = lossp; sl[wave][1] = dcp; }
    blockColumnAtomics<NCH>(dwp, H, dw);   // synchronizes the block before sl is read
    if (threadIdx.x == 0) {
        if (loss) atomicAdd(loss, sl[0][0] + sl[1][0] + sl[2][0] + sl[3][0]);
        atomicAdd(db, sl[0][1] + sl[1][1] + sl[2][1] + sl[3][1]);
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
    const unsigned rowBlocks = (unsigned)std::max(1, std::min(kRowBlocks, (n + 15) / 16));
    switch ((Hd + 63) / 64) {
#define ZG_OUT(C) case C: hipLaunchKernelGGL(outputKernel<C>, dim3(rowBlocks), dim3(256), 0, s, m.act[L - 1], m.pre[L - 1], \
                                             n, Hd, wOut, bOut, y, 1.0f / (float)n, out, m.dz[0], dwOut, dbOut, \
                                             loss, mode); break
        ZG_OUT(1); ZG_OUT(2); ZG_OUT(3); ZG_OUT(4); ZG_OUT(5); ZG_OUT(6); ZG_OUT(7); ZG_OUT(8);
#undef ZG_OUT
        default: return hipErrorInvalidValue;   // hidden is validated to 1..512
    }
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
            g.dbias = m.grad + m.bOffset[l];   // bias gradient of layer l: the row sums of dZ_l^T
            int split = std::max(1, std::min(256, n / 512));
            e = gemm(kBwdWeight, g, split, s);
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
