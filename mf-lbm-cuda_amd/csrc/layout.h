// Host construction of the fluid-compact population layout (see kernels.inl, Sparse<T>).
#pragma once
#include <algorithm>
#include <cstdint>
#include <string>
#include <vector>
#include "lattice.h"

namespace mflbm {

struct LayoutHost {
    int nf = 0;                     // interior fluid nodes
    int64_t ns = 0;                 // cells per slot array (nf + max extra cells of a slot)
    std::vector<int32_t> nb;        // 18 * nf
    // short neighbour rows (Sparse<T>::nb16 / nbb): 9 * nf pairs of 16-bit links, NB_WAVE
    // words per wave of 64 nodes; n_escaped waves keep their 32-bit rows
    std::vector<uint32_t> nb16;
    std::vector<int32_t> nbb;
    int n_escaped = 0;
    std::vector<int32_t> ext_node;  // s1 node of every extra cell, grouped by slot
    std::vector<int32_t> ext_off;   // 20 offsets into ext_node
    std::vector<int32_t> pairs[2];  // periodic copies (slot, dst, src) for [0] even, [1] odd steps
    std::vector<int32_t> zmap;      // 6 planes * 19 slots * X1 * Y1 (open z only)
    std::vector<int32_t> pmap;      // porous plate: 3 planes (zp-1, zp, zp+1) * 19 slots * X1 * Y1
    std::vector<int32_t> zstart;    // nz + 1 offsets of the z-slices in the fluid list
};

// Short neighbour rows (Sparse<T>::nb16 / nbb) of the waves w0 .. w1-1 of 64 nodes, from the
// 32-bit rows nb (18 * nf, cell of node t in direction q at (q - 1) * nf + t; cells >= nf are
// extra cells).  Nodes with !valid(t) (padding) are neither encoded nor bound.  nb16 (9 * nf)
// and nbb (NB_WAVE per wave) are written for these waves only; every link is decoded as the
// kernels decode it and compared.  Returns the number of flagged (32-bit) waves, or -1 when a
// link does not decode.  Disjoint wave ranges may be encoded concurrently.
template <class Valid>
int encode_short_rows(const int32_t* nb, int64_t nf, int64_t w0, int64_t w1, Valid&& valid, uint32_t* nb16, int32_t* nbb) {
    int escaped = 0;
    for (int64_t w = w0; w < w1; ++w) {
        const int64_t t0 = w * 64, t1 = std::min<int64_t>(nf, t0 + 64);
        int32_t* B = nbb + w * NB_WAVE;
        for (int m = 0; m < NB_WAVE; ++m) B[m] = 0;
        for (int64_t t = t0; t < t1; ++t)
            for (int p = 0; p < 9; ++p) nb16[p * nf + t] = 0u;
        bool esc = false;
        for (int q = 1; q < 19; ++q)
            for (int c = 0; c < 2; ++c) {
                int32_t lo = INT32_MAX, hi = INT32_MIN;
                for (int64_t t = t0; t < t1; ++t) {
                    if (!valid(t)) continue;
                    const int32_t v = nb[(q - 1) * nf + t];
                    if ((v >= nf) != (c == 1)) continue;
                    lo = std::min(lo, v);
                    hi = std::max(hi, v);
                }
                if (lo == INT32_MAX) lo = hi = 0;
                if ((int64_t)hi - lo >= 32768) esc = true;
                B[2 * (q - 1) + c] = lo;
            }
        if (esc) {
            B[36] = 1;
            ++escaped;
            continue;
        }
        for (int q = 1; q < 19; ++q)
            for (int64_t t = t0; t < t1; ++t) {
                if (!valid(t)) continue;
                const int32_t v = nb[(q - 1) * nf + t];
                const int c = v >= nf;
                const uint32_t u = (uint32_t)(v - B[2 * (q - 1) + c]) | (uint32_t)c << 15;
                nb16[((q - 1) >> 1) * nf + t] |= u << (16 * ((q - 1) & 1));
            }
        for (int q = 1; q < 19; ++q)
            for (int64_t t = t0; t < t1; ++t) {
                if (!valid(t)) continue;
                const uint32_t pr = nb16[((q - 1) >> 1) * nf + t];
                const uint32_t u = ((q - 1) & 1) ? pr >> 16 : pr & 0xFFFFu;
                if (B[2 * (q - 1) + (int)(u >> 15)] + (int32_t)(u & 0x7FFFu) != nb[(q - 1) * nf + t]) return -1;
            }
    }
    return escaped;
}

// fluid: packed coordinates of the interior fluid nodes in (k, j, i) order.
// in_bc / out_bc: the inlet / outlet condition of an open z axis (1 velocity inlet or
// convective outlet, 2 Zou-He, 0 none); plate_z: plane of the porous plate (0 = none).
bool build_layout(const Dims& d, const std::vector<uint32_t>& fluid, int kper, int jper, int open_z, int in_bc,
                  int out_bc, int plate_z, LayoutHost& out, std::string& err);

}  // namespace mflbm
