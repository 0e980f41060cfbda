// Host construction of the fluid-compact population layout (see kernels.inl, Sparse<T>).
#pragma once
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

// fluid: packed coordinates of the interior fluid nodes in (k, j, i) order.
// in_bc / out_bc: the inlet / outlet condition of an open z axis (1 velocity inlet or
// convective outlet, 2 Zou-He, 0 none); plate_z: plane of the porous plate (0 = none).
bool build_layout(const Dims& d, const std::vector<uint32_t>& fluid, int kper, int jper, int open_z, int in_bc,
                  int out_bc, int plate_z, LayoutHost& out, std::string& err);

}  // namespace mflbm
