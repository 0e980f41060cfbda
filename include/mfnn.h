/*
 * mfnn.h -- C ABI of the warm-start network kernels (gfx950, bf16 MFMA, fp32 accumulate).
 *
 * No counterpart in the reference: its runs start from rest (u = 0,
 * src/Init_multiphase.cpp:569-572).  These kernels train and evaluate the 3-D convolutional
 * network whose predicted flow field is handed to mflbm_init_equilibrium (the hook at
 * src/Init_multiphase.cpp:389-443).  All pointers are device pointers; `stream` is a
 * hipStream_t (null = default stream); every call is asynchronous and returns 0, or < 0 for
 * invalid arguments (-1) or a launch error (-2).
 *
 * Tensors are channels-last [D][H][W][C] (x fastest among the voxels), bf16 unless noted.
 * A 3x3x3 convolution has weights W[n][k], n = output channel (rows padded to 16),
 * k = tap * cin + ci (row length kp, a multiple of 32 >= 27 * cin, zero padded),
 * tap = (dz + 1) * 9 + (dy + 1) * 3 + (dx + 1).
 */
#ifndef MFNN_H
#define MFNN_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Device memory of the network (the same HIP runtime as the kernels): allocate (null on
 * failure), free, synchronous copy (kind 0 host->device, 1 device->host, 2 device->device),
 * zero fill, and a device-wide synchronisation. */
void* mfnn_alloc(int64_t bytes);
int mfnn_free(void* p);
int mfnn_copy(void* dst, const void* src, int64_t bytes, int32_t kind);
int mfnn_zero(void* dst, int64_t bytes);
int mfnn_sync(void);

/* y = conv3x3x3(x, w) + bias, then ReLU (relu != 0) or a multiply by [gate > 0] (gate != null,
 * gate has cout channels); stored to yb (bf16) and/or yf (fp32, either may be null).
 * cin % 8 == 0, cout <= 192 (w has 16 * ceil(cout / 16) rows, or 192 when cout > 128),
 * D, H, W > 0, x and w 16-byte aligned (returns -1 otherwise).
 * With w from mfnn_flip_weights this is the data gradient. */
int mfnn_conv3(const void* x, const void* w, const float* bias, const void* gate, int32_t D, int32_t H, int32_t W,
               int32_t cin, int32_t kp, int32_t cout, int32_t relu, void* yb, float* yf, void* stream);

/* dw[n][k] (+)= sum_voxels dz[m][n] * x[m + off(tap)][ci] and db[n] (+)= sum_m dz[m][n]
 * (accumulate != 0 adds to the existing values).  The voxel sum runs in chunks of `chunk`
 * voxels (multiple of 32) whose fp32 partials go to `work` (mfnn_wgrad_workspace floats).
 * cin and cout are multiples of 8, D, H, W > 0 and x, dz 16-byte aligned (returns -1
 * otherwise). */
int64_t mfnn_wgrad_workspace(int32_t D, int32_t H, int32_t W, int32_t kp, int32_t cout, int32_t chunk);
int mfnn_conv3_wgrad(const void* x, const void* dz, int32_t D, int32_t H, int32_t W, int32_t cin, int32_t kp,
                     int32_t cout, int32_t chunk, float* work, float* dw, float* db, int32_t accumulate, void* stream);

/* mfnn_conv3_wgrad without the workspace argument, for callers that cannot allocate device
 * memory (code under HIP graph capture, tight training loops, harnesses that forbid
 * hipMalloc).  No counterpart in the reference.  The same kernels write their fp32 partials
 * to a fixed array of MFNN_WGRAD_SCRATCH_FLOATS floats (64 MiB) that the library holds on
 * every device it runs on until the process ends; the call allocates nothing.  Returns -3
 * (and launches nothing) when mfnn_wgrad_workspace(D, H, W, kp, cout, chunk) exceeds that
 * size, otherwise as mfnn_conv3_wgrad, with results bit-identical to it.  Calls on one
 * device share its array: calls that overlap on one device (other streams, other host
 * threads) race and must be ordered by the caller. */
#define MFNN_WGRAD_SCRATCH_FLOATS (16 * 1024 * 1024)
int mfnn_conv3_wgrad_static(const void* x, const void* dz, int32_t D, int32_t H, int32_t W, int32_t cin, int32_t kp,
                            int32_t cout, int32_t chunk, float* dw, float* db, int32_t accumulate, void* stream);

/* wt[ci][tap' * cout + co] = w[co][(26 - tap') * cin + ci] (rows padded to cinp, a multiple of
 * 16; row length kp2, a multiple of 32 >= 27 * cout): the weights of the data gradient. */
int mfnn_flip_weights(const void* w, int32_t cin, int32_t kp, int32_t cout, int32_t kp2, int32_t cinp, void* wt,
                      void* stream);

/* fp32 -> bf16 copy of n values */
int mfnn_to_bf16(const float* x, int64_t n, void* y, void* stream);

/* 2x2x2 average pooling of a D x H x W grid (even sizes) and, backward != 0, its adjoint
 * (x = gradient of the pooled grid, y = gradient of the fine grid). */
int mfnn_pool2(const void* x, int32_t D, int32_t H, int32_t W, int32_t C, void* y, int32_t backward, void* stream);

/* nearest 2x upsampling of a coarse grid into channels [c0, c0 + C) of a ct-channel
 * D x H x W tensor and, backward != 0, its adjoint (fine gradient -> coarse gradient). */
int mfnn_up2(const void* x, int32_t D, int32_t H, int32_t W, int32_t C, int32_t ct, int32_t c0, void* y,
             int32_t backward, void* stream);

/* copy channels [c0, c0 + C) between a C-channel and a ct-channel tensor of nvox voxels
 * (into_wide != 0: narrow -> wide, else wide -> narrow). */
int mfnn_channels(const void* src, int64_t nvox, int32_t C, int32_t ct, int32_t c0, int32_t into_wide, void* dst,
                  void* stream);

/* out = (a + b) * [gate > 0] over n values (b, gate may be null). */
int mfnn_add_gate(const void* a, const void* b, const void* gate, int64_t n, void* out, void* stream);

/* masked squared error of fp32 y against t (M voxels x cout channels, mask per voxel):
 * part[block] = sum of (mask (y - t))^2 over the block's values (fp64, mfnn_mse_partials
 * blocks), dy = 2 mask (y - t) * inv_count (bf16). */
int64_t mfnn_mse_partials(int64_t M, int32_t cout);
int mfnn_mse(const float* y, const float* t, const float* mask, int64_t M, int32_t cout, float inv_count, void* dy,
             double* part, void* stream);

/* Adam step (bias-corrected for `step` >= 1) on fp32 weights w with gradient g and moments
 * m1, m2; writes the bf16 copy wb when not null. */
int mfnn_adam(float* w, const float* g, float* m1, float* m2, int64_t n, float lr, float b1, float b2, float eps,
              int32_t step, void* wb, void* stream);

#ifdef __cplusplus
}
#endif

#endif
