// Kernels of the learned warm start: a 3-D convolutional U-Net that maps a pore geometry
// (solid mask + signed distance) to an initial flow field.  No counterpart in the reference
// (its initial state is u = 0, src/Init_multiphase.cpp:569-572); the network output feeds
// the equilibrium initialisation (mflbm_init_equilibrium).
//
// Tensors are channels-last: [D][H][W][C], C a multiple of 8 for bf16 operands so that one
// lane fetches 8 consecutive channels in one 16-byte load.  Every convolution is 3x3x3,
// stride 1, zero padding 1, computed as an implicit GEMM on the MFMA units with bf16
// operands and fp32 accumulation (v_mfma_f32_16x16x32_bf16):
//   forward   Y[m][n] = sum_k A[m][k] B[k][n],  m = voxel, n = output channel,
//             k = tap * Cin + ci, A[m][k] = X[m + off(tap)][ci], B[k][n] = W[n][k];
//   data grad the same kernel on dY with the weights re-laid out as W'[ci][tap' * Cout + co]
//             = W[co][(26 - tap') * Cin + ci] (off(26 - t) = -off(t));
//   weight grad dW[n][k] = sum_m dZ[m][n] A[m][k], the voxel sum split into chunks whose fp32
//             partial tiles are added in chunk order (deterministic).
// Weights are kept as W[n][k] with n padded to 16 and k to 32 (zero padding), fp32 master
// copy plus a bf16 copy for the MFMA operands.
#include <hip/hip_runtime.h>
#include <cstdint>
#include <mutex>

#include "../../include/mfnn.h"
#include "device_lock.h"

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ float bf2f(__bf16 v) { return (float)v; }

struct Grid3 {
    int D, H, W;
};

// Taps: 0..26, x fastest (dx = tap % 3 - 1, dy = tap / 3 % 3 - 1, dz = tap / 9 - 1); a tap
// outside the grid reads zero (zero padding).

// Forward / data-gradient convolution.  One wave: MT x 16 voxels x NT*16 output channels,
// the 16-channel tiles blockIdx.y * NT .. + NT - 1 (small grids split the channels over
// blocks); the MT voxel tiles share each weight load.
// Epilogue: + bias (if given), then either ReLU, or a multiply by the ReLU mask of `gate`
// (the data gradient of a layer whose input came through a ReLU), stored as bf16 and/or fp32.
template <int NT, int MT>
__global__ __launch_bounds__(256) void k_conv3(const __bf16* __restrict__ X, const __bf16* __restrict__ Wb,
                                               const float* __restrict__ bias, const __bf16* __restrict__ gate,
                                               Grid3 g, int cin, int kp, int cout, int relu,
                                               __bf16* __restrict__ yb, float* __restrict__ yf) {
    const int M = g.D * g.H * g.W;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int m0 = (blockIdx.x * 4 + wave) * 16 * MT;
    if (m0 >= M) return;
    const int r = lane & 15, q = lane >> 4;
    const int K = 27 * cin;
    const int tb = blockIdx.y * NT;  // first channel tile of this block
    // this lane's voxel of each voxel tile and its coordinates (fixed over the k loop)
    int mv[MT], xv[MT], yv[MT], zv[MT];
#pragma unroll
    for (int u = 0; u < MT; ++u) {
        mv[u] = m0 + 16 * u + r;
        xv[u] = mv[u] % g.W;
        yv[u] = (mv[u] / g.W) % g.H;
        zv[u] = mv[u] / (g.W * g.H);
    }
    f32x4 acc[MT][NT];
#pragma unroll
    for (int u = 0; u < MT; ++u)
#pragma unroll
        for (int t = 0; t < NT; ++t) acc[u][t] = f32x4{0.f, 0.f, 0.f, 0.f};
    for (int kb = 0; kb < kp; kb += 32) {
        const int k = kb + 8 * q;
        const int tap = k < K ? k / cin : 0, ci = k - tap * cin;
        const int ox = tap % 3 - 1, oy = (tap / 3) % 3 - 1, oz = tap / 9 - 1;
        bf16x8 b[NT];
#pragma unroll
        for (int t = 0; t < NT; ++t)
            b[t] = *reinterpret_cast<const bf16x8*>(Wb + (int64_t)((tb + t) * 16 + r) * kp + k);
#pragma unroll
        for (int u = 0; u < MT; ++u) {
            bf16x8 a;
#pragma unroll
            for (int j = 0; j < 8; ++j) a[j] = (__bf16)0.f;
            const int xx = xv[u] + ox, yy = yv[u] + oy, zz = zv[u] + oz;
            if (k < K && mv[u] < M && xx >= 0 && yy >= 0 && zz >= 0 && xx < g.W && yy < g.H && zz < g.D)
                a = *reinterpret_cast<const bf16x8*>(X + (int64_t)((zz * g.H + yy) * g.W + xx) * cin + ci);
#pragma unroll
            for (int t = 0; t < NT; ++t)
                acc[u][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b[t], acc[u][t], 0, 0, 0);
        }
    }
#pragma unroll
    for (int u = 0; u < MT; ++u)
#pragma unroll
        for (int t = 0; t < NT; ++t) {
            const int n = (tb + t) * 16 + r;
            if (n >= cout) continue;
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int mm = m0 + 16 * u + 4 * q + j;
                if (mm >= M) continue;
                float v = acc[u][t][j] + (bias ? bias[n] : 0.f);
                if (relu) v = v > 0.f ? v : 0.f;
                if (gate && !(bf2f(gate[(int64_t)mm * cout + n]) > 0.f)) v = 0.f;
                if (yb) yb[(int64_t)mm * cout + n] = (__bf16)v;
                if (yf) yf[(int64_t)mm * cout + n] = v;
            }
        }
}

// Weight gradient partials.  Block (kt, nt, chunk): one wave per 16 x 16·KT block of (output
// channel n, k) over `chunk` voxels, as KT 16 x 16 MFMA tiles that share the dZ operand;
// writes part[chunk][n][k] (fp32).  The waves of the first k block also sum the dZ values
// they load: the bias gradient partials bpart[chunk][n] (n < cout; null: not wanted).
// Per step of 32 voxels the operands are staged through LDS: lane (v, h) loads 8 channels
// (16 bytes) of voxel v -- dZ channels n0 + 8h .., and of each tile the 8 columns
// k0 + 16t + 8h .. (8 consecutive channels of one tap, since cin is a multiple of 8) at the
// tap's shifted voxel -- and lane (r, q) then reads column r of voxels 8q .. 8q + 7, the
// MFMA operand layout.
template <int KT>
__global__ __launch_bounds__(64) void k_conv3_wgrad(const __bf16* __restrict__ X, const __bf16* __restrict__ dZ,
                                                    Grid3 g, int cin, int kp, int coutp, int cout, int chunk,
                                                    float* __restrict__ part, float* __restrict__ bpart) {
    __shared__ __attribute__((aligned(16))) __bf16 sA[32][16];
    __shared__ __attribute__((aligned(16))) __bf16 sB[KT][32][16];
    const int M = g.D * g.H * g.W;
    const int lane = threadIdx.x;
    const int k0 = blockIdx.x * 16 * KT, n0 = blockIdx.y * 16, c = blockIdx.z;
    const int r = lane & 15, q = lane >> 4;  // MFMA role: column r, voxels 8q .. 8q + 7
    const int v = lane >> 1, h = lane & 1;   // loader role: voxel v, 8-channel half h
    const int K = 27 * cin;
    const int n = n0 + r;
    const bool ain = n0 + 8 * h < cout;  // cout is a multiple of 8: the whole half or none
    int lci[KT], lox[KT], loy[KT], loz[KT], loff[KT];
    bool lkin[KT];
#pragma unroll
    for (int t = 0; t < KT; ++t) {
        const int kk = k0 + 16 * t + 8 * h;
        lkin[t] = kk < K;  // K is a multiple of 8: the whole half or none
        const int tap = lkin[t] ? kk / cin : 0;
        lci[t] = lkin[t] ? kk - tap * cin : 0;
        lox[t] = tap % 3 - 1;
        loy[t] = (tap / 3) % 3 - 1;
        loz[t] = tap / 9 - 1;
        loff[t] = (loz[t] * g.H + loy[t]) * g.W + lox[t];
    }
    f32x4 acc[KT];
#pragma unroll
    for (int t = 0; t < KT; ++t) acc[t] = f32x4{0.f, 0.f, 0.f, 0.f};
    float bsum = 0.f;  // this lane's dZ values (row n), in voxel order
    const int mb = c * chunk, me = min(M, mb + chunk);
    bf16x8 zero;
#pragma unroll
    for (int j = 0; j < 8; ++j) zero[j] = (__bf16)0.f;
    for (int ms = mb; ms < me; ms += 32) {
        const int m = ms + v;
        bf16x8 la = zero;
        if (m < me && ain) la = *reinterpret_cast<const bf16x8*>(dZ + (int64_t)m * cout + n0 + 8 * h);
        *reinterpret_cast<bf16x8*>(&sA[v][8 * h]) = la;
        const int x = m % g.W, y = (m / g.W) % g.H, z = m / (g.W * g.H);
#pragma unroll
        for (int t = 0; t < KT; ++t) {
            const int xx = x + lox[t], yy = y + loy[t], zz = z + loz[t];
            const bool in = m < me && lkin[t] && xx >= 0 && yy >= 0 && zz >= 0 && xx < g.W && yy < g.H && zz < g.D;
            bf16x8 lb = zero;
            if (in) lb = *reinterpret_cast<const bf16x8*>(X + (int64_t)(m + loff[t]) * cin + lci[t]);
            *reinterpret_cast<bf16x8*>(&sB[t][v][8 * h]) = lb;
        }
        __syncthreads();
        bf16x8 a, b[KT];
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            a[j] = sA[8 * q + j][r];
#pragma unroll
            for (int t = 0; t < KT; ++t) b[t][j] = sB[t][8 * q + j][r];
        }
        if (bpart && k0 == 0) {
#pragma unroll
            for (int j = 0; j < 8; ++j) bsum += bf2f(a[j]);
        }
#pragma unroll
        for (int t = 0; t < KT; ++t) acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b[t], acc[t], 0, 0, 0);
        __syncthreads();
    }
    if (bpart && k0 == 0) {
        // lanes r, r + 16, r + 32, r + 48 hold row n: fixed pairing order
        const float s1 = bsum + __shfl_xor(bsum, 16);
        const float s2 = s1 + __shfl_xor(s1, 32);
        if (q == 0 && n < cout) bpart[(int64_t)c * cout + n] = s2;
    }
    // tile t: D[row = n0 + 4q + j][col = k0 + 16t + r]
#pragma unroll
    for (int t = 0; t < KT; ++t)
#pragma unroll
        for (int j = 0; j < 4; ++j) part[((int64_t)c * coutp + n0 + 4 * q + j) * kp + k0 + 16 * t + r] = acc[t][j];
}

// dW[i] (+)= sum over chunks of part[chunk][i], i < n, and in the blocks after those,
// db[i] (+)= the same over bpart[chunk][i], i < nb (bpart null: no bias).  A block covers
// 32 entries with 8 groups of 32 threads: group g adds chunks g, g + 8, g + 16, ... in order,
// then the 8 group sums are added in group order (a fixed order for every launch).
__global__ __launch_bounds__(256) void k_sum_chunks(const float* __restrict__ part, int nchunk, int64_t n,
                                                    const float* __restrict__ bpart, int nb, int accumulate,
                                                    float* __restrict__ dw, float* __restrict__ db) {
    __shared__ float red[8][32];
    const int x = threadIdx.x & 31, grp = threadIdx.x >> 5;
    const int64_t wblocks = (n + 31) / 32;
    const bool bias = (int64_t)blockIdx.x >= wblocks;
    const float* src = bias ? bpart : part;
    float* dst = bias ? db : dw;
    const int64_t len = bias ? nb : n;
    const int64_t i = ((int64_t)blockIdx.x - (bias ? wblocks : 0)) * 32 + x;
    float s = 0.f;
    if (i < len)
        for (int c = grp; c < nchunk; c += 8) s += src[(int64_t)c * len + i];
    red[grp][x] = s;
    __syncthreads();
    if (grp == 0 && i < len) {
        float t = red[0][x];
#pragma unroll
        for (int g = 1; g < 8; ++g) t += red[g][x];
        dst[i] = accumulate ? dst[i] + t : t;
    }
}

// W'[ci][tap' * cout + co] = W[co][(26 - tap') * cin + ci]  (bf16, padded: ci to 16, k' to 32)
__global__ void k_flip_weights(const __bf16* __restrict__ W, int cin, int kp, int cout, int kp2, int cinp,
                               __bf16* __restrict__ Wt) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int64_t)cinp * kp2) return;
    const int ci = (int)(i / kp2), k2 = (int)(i % kp2);
    __bf16 v = (__bf16)0.f;
    if (ci < cin && k2 < 27 * cout) {
        const int tp = k2 / cout, co = k2 - tp * cout;
        v = W[(int64_t)co * kp + (26 - tp) * cin + ci];
    }
    Wt[i] = v;
}

__global__ void k_to_bf16(const float* __restrict__ x, int64_t n, __bf16* __restrict__ y) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = (__bf16)x[i];
}

// 2x2x2 average pool (fwd) and its adjoint; nearest 2x upsampling (fwd) and its adjoint.
// Fine grid D x H x W (even), coarse D/2 x H/2 x W/2, C channels.
__global__ void k_pool2(const __bf16* __restrict__ x, Grid3 f, int C, __bf16* __restrict__ y) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int Wc = f.W / 2, Hc = f.H / 2, Dc = f.D / 2;
    if (i >= (int64_t)Dc * Hc * Wc * C) return;
    const int c = (int)(i % C);
    const int64_t v = i / C;
    const int xc = (int)(v % Wc), yc = (int)((v / Wc) % Hc), zc = (int)(v / ((int64_t)Wc * Hc));
    float s = 0.f;
    for (int dz = 0; dz < 2; ++dz)
        for (int dy = 0; dy < 2; ++dy)
            for (int dx = 0; dx < 2; ++dx)
                s += bf2f(x[(((int64_t)(2 * zc + dz) * f.H + 2 * yc + dy) * f.W + 2 * xc + dx) * C + c]);
    y[i] = (__bf16)(0.125f * s);
}

__global__ void k_pool2_bwd(const __bf16* __restrict__ dy, Grid3 f, int C, __bf16* __restrict__ dx) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int64_t)f.D * f.H * f.W * C) return;
    const int c = (int)(i % C);
    const int64_t v = i / C;
    const int x = (int)(v % f.W), y = (int)((v / f.W) % f.H), z = (int)(v / ((int64_t)f.W * f.H));
    const int Wc = f.W / 2, Hc = f.H / 2;
    dx[i] = (__bf16)(0.125f * bf2f(dy[(((int64_t)(z / 2) * Hc + y / 2) * Wc + x / 2) * C + c]));
}

// upsample from coarse into channels [c0, c0 + C) of a fine tensor with Ct channels
__global__ void k_up2(const __bf16* __restrict__ x, Grid3 f, int C, int Ct, int c0, __bf16* __restrict__ y) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int64_t)f.D * f.H * f.W * C) return;
    const int c = (int)(i % C);
    const int64_t v = i / C;
    const int xx = (int)(v % f.W), yy = (int)((v / f.W) % f.H), zz = (int)(v / ((int64_t)f.W * f.H));
    const int Wc = f.W / 2, Hc = f.H / 2;
    y[v * Ct + c0 + c] = x[(((int64_t)(zz / 2) * Hc + yy / 2) * Wc + xx / 2) * C + c];
}

__global__ void k_up2_bwd(const __bf16* __restrict__ dy, Grid3 f, int C, int Ct, int c0, __bf16* __restrict__ dx) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int Wc = f.W / 2, Hc = f.H / 2, Dc = f.D / 2;
    if (i >= (int64_t)Dc * Hc * Wc * C) return;
    const int c = (int)(i % C);
    const int64_t v = i / C;
    const int xc = (int)(v % Wc), yc = (int)((v / Wc) % Hc), zc = (int)(v / ((int64_t)Wc * Hc));
    float s = 0.f;
    for (int dz = 0; dz < 2; ++dz)
        for (int dy2 = 0; dy2 < 2; ++dy2)
            for (int dx2 = 0; dx2 < 2; ++dx2)
                s += bf2f(dy[((((int64_t)(2 * zc + dz) * f.H + 2 * yc + dy2) * f.W) + 2 * xc + dx2) * Ct + c0 + c]);
    dx[i] = (__bf16)s;
}

// copy channels [c0, c0 + C) between a C-channel tensor and a Ct-channel tensor
__global__ void k_channels(const __bf16* __restrict__ src, int64_t nvox, int C, int Ct, int c0, int into_wide,
                           __bf16* __restrict__ dst) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nvox * C) return;
    const int64_t v = i / C;
    const int c = (int)(i % C);
    if (into_wide) dst[v * Ct + c0 + c] = src[i];
    else dst[i] = src[v * Ct + c0 + c];
}

// out = (a + b) * [gate > 0]  (b and gate optional): gradients meeting at a branch point and
// the ReLU derivative
__global__ void k_add_gate(const __bf16* __restrict__ a, const __bf16* __restrict__ b, const __bf16* __restrict__ gate,
                           int64_t n, __bf16* __restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float v = bf2f(a[i]) + (b ? bf2f(b[i]) : 0.f);
    if (gate && !(bf2f(gate[i]) > 0.f)) v = 0.f;
    out[i] = (__bf16)v;
}

// masked mean-squared error over the first `cout` channels: loss partials per block and
// dY = 2 (y - t) mask / count as bf16
__global__ __launch_bounds__(256) void k_mse(const float* __restrict__ y, const float* __restrict__ t,
                                             const float* __restrict__ mask, int64_t M, int cout, float inv_count,
                                             __bf16* __restrict__ dy, double* __restrict__ part) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    double e = 0.0;
    if (i < M * cout) {
        const float w = mask[i / cout];
        const float d = (y[i] - t[i]) * w;
        e = (double)d * d;
        dy[i] = (__bf16)(2.f * d * inv_count);
    }
    __shared__ double red[256];
    red[threadIdx.x] = e;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// Adam step on fp32 master weights; refreshes the bf16 copy.
__global__ void k_adam(float* __restrict__ w, const float* __restrict__ g, float* __restrict__ m1,
                       float* __restrict__ m2, int64_t n, float lr, float b1, float b2, float eps, float c1,
                       float c2, __bf16* __restrict__ wb) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float a = b1 * m1[i] + (1.f - b1) * gi;
    const float b = b2 * m2[i] + (1.f - b2) * gi * gi;
    m1[i] = a;
    m2[i] = b;
    const float v = w[i] - lr * (a * c1) / (sqrtf(b * c2) + eps);
    w[i] = v;
    if (wb) wb[i] = (__bf16)v;
}

inline unsigned nblk(int64_t n, int b = 256) { return (unsigned)((n + b - 1) / b); }

}  // namespace

extern "C" {

void* mfnn_alloc(int64_t bytes) {
    std::lock_guard<std::recursive_mutex> lock(mflbm::device_lock());
    void* p = nullptr;
    if (bytes <= 0 || hipMalloc(&p, (size_t)bytes) != hipSuccess) return nullptr;
    return p;
}

int mfnn_free(void* p) {
    std::lock_guard<std::recursive_mutex> lock(mflbm::device_lock());
    return hipFree(p) == hipSuccess ? 0 : -2;
}

int mfnn_copy(void* dst, const void* src, int64_t bytes, int32_t kind) {
    const hipMemcpyKind k = kind == 0 ? hipMemcpyHostToDevice : (kind == 1 ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice);
    if (bytes < 0 || kind < 0 || kind > 2) return -1;
    return hipMemcpy(dst, src, (size_t)bytes, k) == hipSuccess ? 0 : -2;
}

int mfnn_zero(void* dst, int64_t bytes) {
    if (bytes < 0) return -1;
    return hipMemset(dst, 0, (size_t)bytes) == hipSuccess ? 0 : -2;
}

int mfnn_sync(void) { return hipDeviceSynchronize() == hipSuccess ? 0 : -2; }

int mfnn_conv3(const void* x, const void* w, const float* bias, const void* gate, int32_t D, int32_t H, int32_t W,
               int32_t cin, int32_t kp, int32_t cout, int32_t relu, void* yb, float* yf, void* stream) {
    if (cin % 8 || kp % 32 || kp < 27 * cin || cout <= 0 || cout > 192 || D <= 0 || H <= 0 || W <= 0) return -1;
    if ((uintptr_t)x % 16 || (uintptr_t)w % 16) return -1;  // 16-byte operand loads
    const Grid3 g{D, H, W};
    const int64_t M = (int64_t)D * H * W;
    const unsigned vblocks = nblk(M, 64);
    hipStream_t s = (hipStream_t)stream;
    // channel tiles the weight rows cover: ceil(cout / 16), or 12 (192 rows) above 128 channels
    const int ntw = cout > 128 ? 12 : (cout + 15) / 16;
    const __bf16* X = (const __bf16*)x;
    const __bf16* Wb = (const __bf16*)w;
    const __bf16* G = (const __bf16*)gate;
    __bf16* Y = (__bf16*)yb;
    if (vblocks < 256) {
        // fewer voxel blocks than CUs: one channel tile per block, the tiles over grid.y
        k_conv3<1, 1><<<dim3(vblocks, ntw), 256, 0, s>>>(X, Wb, bias, G, g, cin, kp, cout, relu, Y, yf);
        return hipGetLastError() == hipSuccess ? 0 : -2;
    }
    const unsigned grid = vblocks;
    switch (ntw) {
        // one or two channel tiles on grids of 65536 voxels and more: two voxel tiles per wave
        // share each weight load (on smaller grids the halved block count costs more)
        case 1:
            if (M >= 65536) k_conv3<1, 2><<<nblk(M, 128), 256, 0, s>>>(X, Wb, bias, G, g, cin, kp, cout, relu, Y, yf);
            else k_conv3<1, 1><<<grid, 256, 0, s>>>(X, Wb, bias, G, g, cin, kp, cout, relu, Y, yf);
            break;
        case 2:
            if (M >= 65536) k_conv3<2, 2><<<nblk(M, 128), 256, 0, s>>>(X, Wb, bias, G, g, cin, kp, cout, relu, Y, yf);
            else k_conv3<2, 1><<<grid, 256, 0, s>>>(X, Wb, bias, G, g, cin, kp, cout, relu, Y, yf);
            break;
        case 3: k_conv3<3, 1><<<grid, 256, 0, s>>>(X, Wb, bias, G, g, cin, kp, cout, relu, Y, yf); break;
        case 4: k_conv3<4, 1><<<grid, 256, 0, s>>>(X, Wb, bias, G, g, cin, kp, cout, relu, Y, yf); break;
        case 5: k_conv3<5, 1><<<grid, 256, 0, s>>>(X, Wb, bias, G, g, cin, kp, cout, relu, Y, yf); break;
        case 6: k_conv3<6, 1><<<grid, 256, 0, s>>>(X, Wb, bias, G, g, cin, kp, cout, relu, Y, yf); break;
        case 7: k_conv3<7, 1><<<grid, 256, 0, s>>>(X, Wb, bias, G, g, cin, kp, cout, relu, Y, yf); break;
        case 8: k_conv3<8, 1><<<grid, 256, 0, s>>>(X, Wb, bias, G, g, cin, kp, cout, relu, Y, yf); break;
        default: k_conv3<12, 1><<<grid, 256, 0, s>>>(X, Wb, bias, G, g, cin, kp, cout, relu, Y, yf); break;
    }
    return hipGetLastError() == hipSuccess ? 0 : -2;
}

int64_t mfnn_wgrad_workspace(int32_t D, int32_t H, int32_t W, int32_t kp, int32_t cout, int32_t chunk) {
    const int64_t M = (int64_t)D * H * W;
    const int64_t nchunk = (M + chunk - 1) / chunk;
    return nchunk * ((int64_t)((cout + 15) / 16 * 16) * kp + cout);  // weight partials, then bias partials
}

}  // extern "C"

namespace {

bool wgrad_args_ok(const void* x, const void* dz, int32_t D, int32_t H, int32_t W, int32_t cin, int32_t kp,
                   int32_t cout, int32_t chunk) {
    if (cin % 8 || cout % 8 || cin <= 0 || cout <= 0 || kp % 32 || kp < 27 * cin || chunk <= 0 || chunk % 32) return false;
    if (D <= 0 || H <= 0 || W <= 0) return false;
    return (uintptr_t)x % 16 == 0 && (uintptr_t)dz % 16 == 0;  // 16-byte operand loads
}

// The weight-gradient launches of valid arguments: partials into `work`, then the chunk sums.
int wgrad_launch(const void* x, const void* dz, int32_t D, int32_t H, int32_t W, int32_t cin, int32_t kp, int32_t cout,
                 int32_t chunk, float* work, float* dw, float* db, int32_t accumulate, void* stream) {
    const Grid3 g{D, H, W};
    const int64_t M = (int64_t)D * H * W;
    const int nchunk = (int)((M + chunk - 1) / chunk);
    const int coutp = (cout + 15) / 16 * 16;
    hipStream_t s = (hipStream_t)stream;
    const int64_t n = (int64_t)coutp * kp;
    float* bpart = db ? work + nchunk * n : nullptr;
    // two k tiles per wave halve the dZ loads, one keeps twice the waves: the second wins
    // when the single-tile grid has fewer than 3072 waves (three per SIMD)
    if ((int64_t)(kp / 16) * (coutp / 16) * nchunk >= 3072)  // kp is a multiple of 32
        k_conv3_wgrad<2><<<dim3(kp / 32, coutp / 16, nchunk), 64, 0, s>>>((const __bf16*)x, (const __bf16*)dz, g,
                                                                          cin, kp, coutp, cout, chunk, work, bpart);
    else
        k_conv3_wgrad<1><<<dim3(kp / 16, coutp / 16, nchunk), 64, 0, s>>>((const __bf16*)x, (const __bf16*)dz, g,
                                                                          cin, kp, coutp, cout, chunk, work, bpart);
    const unsigned sblocks = nblk(n, 32) + (db ? nblk(cout, 32) : 0);
    k_sum_chunks<<<sblocks, 256, 0, s>>>(work, nchunk, n, bpart, db ? cout : 0, accumulate, dw, db);
    return hipGetLastError() == hipSuccess ? 0 : -2;
}

// Scratch of mfnn_conv3_wgrad_static: every partial it sums is written first by the
// partial kernels of the same call, so no call needs it cleared.  In host code `scratch`
// names a host-side shadow, not the device array: the device address comes from
// hipGetSymbolAddress, once per device (each device holds its own copy of the array).
static __device__ __align__(256) float scratch[MFNN_WGRAD_SCRATCH_FLOATS];
constexpr int MAX_DEVICES = 64;
std::mutex scratch_mutex;
float* scratch_dev[MAX_DEVICES] = {};

// Address of the current device's scratch (null on a runtime error).
float* scratch_address() {
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= MAX_DEVICES) return nullptr;
    std::lock_guard<std::mutex> lock(scratch_mutex);
    if (!scratch_dev[dev]) {
        void* p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess) return nullptr;
        scratch_dev[dev] = static_cast<float*>(p);
    }
    return scratch_dev[dev];
}

}  // namespace

extern "C" {

int mfnn_conv3_wgrad(const void* x, const void* dz, int32_t D, int32_t H, int32_t W, int32_t cin, int32_t kp,
                     int32_t cout, int32_t chunk, float* work, float* dw, float* db, int32_t accumulate, void* stream) {
    if (!wgrad_args_ok(x, dz, D, H, W, cin, kp, cout, chunk)) return -1;
    return wgrad_launch(x, dz, D, H, W, cin, kp, cout, chunk, work, dw, db, accumulate, stream);
}

int mfnn_conv3_wgrad_static(const void* x, const void* dz, int32_t D, int32_t H, int32_t W, int32_t cin, int32_t kp,
                            int32_t cout, int32_t chunk, float* dw, float* db, int32_t accumulate, void* stream) {
    if (!wgrad_args_ok(x, dz, D, H, W, cin, kp, cout, chunk)) return -1;
    if (mfnn_wgrad_workspace(D, H, W, kp, cout, chunk) > (int64_t)MFNN_WGRAD_SCRATCH_FLOATS) return -3;
    float* work = scratch_address();
    if (!work) return -2;
    return wgrad_launch(x, dz, D, H, W, cin, kp, cout, chunk, work, dw, db, accumulate, stream);
}

int mfnn_flip_weights(const void* w, int32_t cin, int32_t kp, int32_t cout, int32_t kp2, int32_t cinp, void* wt,
                      void* stream) {
    if (kp2 < 27 * cout || kp2 % 32 || cinp < cin || cinp % 16) return -1;
    const int64_t n = (int64_t)cinp * kp2;
    k_flip_weights<<<nblk(n), 256, 0, (hipStream_t)stream>>>((const __bf16*)w, cin, kp, cout, kp2, cinp, (__bf16*)wt);
    return hipGetLastError() == hipSuccess ? 0 : -2;
}

int mfnn_to_bf16(const float* x, int64_t n, void* y, void* stream) {
    k_to_bf16<<<nblk(n), 256, 0, (hipStream_t)stream>>>(x, n, (__bf16*)y);
    return hipGetLastError() == hipSuccess ? 0 : -2;
}

int mfnn_pool2(const void* x, int32_t D, int32_t H, int32_t W, int32_t C, void* y, int32_t backward, void* stream) {
    if (D % 2 || H % 2 || W % 2) return -1;
    const Grid3 f{D, H, W};
    hipStream_t s = (hipStream_t)stream;
    if (!backward) k_pool2<<<nblk((int64_t)D * H * W * C / 8), 256, 0, s>>>((const __bf16*)x, f, C, (__bf16*)y);
    else k_pool2_bwd<<<nblk((int64_t)D * H * W * C), 256, 0, s>>>((const __bf16*)x, f, C, (__bf16*)y);
    return hipGetLastError() == hipSuccess ? 0 : -2;
}

int mfnn_up2(const void* x, int32_t D, int32_t H, int32_t W, int32_t C, int32_t ct, int32_t c0, void* y,
             int32_t backward, void* stream) {
    if (D % 2 || H % 2 || W % 2 || c0 < 0 || c0 + C > ct) return -1;
    const Grid3 f{D, H, W};
    hipStream_t s = (hipStream_t)stream;
    if (!backward) k_up2<<<nblk((int64_t)D * H * W * C), 256, 0, s>>>((const __bf16*)x, f, C, ct, c0, (__bf16*)y);
    else k_up2_bwd<<<nblk((int64_t)D * H * W * C / 8), 256, 0, s>>>((const __bf16*)x, f, C, ct, c0, (__bf16*)y);
    return hipGetLastError() == hipSuccess ? 0 : -2;
}

int mfnn_channels(const void* src, int64_t nvox, int32_t C, int32_t ct, int32_t c0, int32_t into_wide, void* dst,
                  void* stream) {
    if (c0 < 0 || c0 + C > ct) return -1;
    k_channels<<<nblk(nvox * C), 256, 0, (hipStream_t)stream>>>((const __bf16*)src, nvox, C, ct, c0, into_wide,
                                                              (__bf16*)dst);
    return hipGetLastError() == hipSuccess ? 0 : -2;
}

int mfnn_add_gate(const void* a, const void* b, const void* gate, int64_t n, void* out, void* stream) {
    k_add_gate<<<nblk(n), 256, 0, (hipStream_t)stream>>>((const __bf16*)a, (const __bf16*)b, (const __bf16*)gate, n,
                                                        (__bf16*)out);
    return hipGetLastError() == hipSuccess ? 0 : -2;
}

int64_t mfnn_mse_partials(int64_t M, int32_t cout) { return (M * cout + 255) / 256; }

int mfnn_mse(const float* y, const float* t, const float* mask, int64_t M, int32_t cout, float inv_count, void* dy,
             double* part, void* stream) {
    k_mse<<<nblk(M * cout), 256, 0, (hipStream_t)stream>>>(y, t, mask, M, cout, inv_count, (__bf16*)dy, part);
    return hipGetLastError() == hipSuccess ? 0 : -2;
}

int mfnn_adam(float* w, const float* g, float* m1, float* m2, int64_t n, float lr, float b1, float b2, float eps,
              int32_t step, void* wb, void* stream) {
    if (step < 1) return -1;
    const float c1 = 1.f / (1.f - powf(b1, (float)step));
    const float c2 = 1.f / (1.f - powf(b2, (float)step));
    k_adam<<<nblk(n), 256, 0, (hipStream_t)stream>>>(w, g, m1, m2, n, lr, b1, b2, eps, c1, c2, (__bf16*)wb);
    return hipGetLastError() == hipSuccess ? 0 : -2;
}

}  // extern "C"
