// Lattice tables, index arithmetic and device-state descriptors shared by the HIP kernels
// and the host side of libmflbm.  D3Q19 velocity set and weights as in
// includes/Module.h:98-122 of the reference; ghost-aware 1-based indexing as in
// includes/Idx_gpu.cuh:49-70.
#pragma once
#include <cstdint>

#define MFLBM_HD __host__ __device__ __forceinline__

namespace mflbm {

constexpr int EX[19] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0};
constexpr int EY[19] = {0, 0, 0, 1, -1, 0, 0, 1, 1, -1, -1, 0, 0, 0, 0, 1, -1, 1, -1};
constexpr int EZ[19] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, 1, -1, -1, 1, 1, -1, -1};
constexpr int OPP[19] = {0, 2, 1, 4, 3, 6, 5, 10, 9, 8, 7, 14, 13, 12, 11, 18, 17, 16, 15};

// Grid extents.  All node indices fit in 32 bits for single domains up to ~1000^3;
// slot offsets (q * S1) are formed in 64 bits.
struct Dims {
    int nx, ny, nz;
    int X1, Y1, Z1;  // nx+2 ...
    int X2, Y2, Z2;  // nx+4 ...
    int X4, Y4, Z4;  // nx+8 ...
    int64_t S1, S2, S4;
};

inline Dims make_dims(int nx, int ny, int nz) {
    Dims d;
    d.nx = nx; d.ny = ny; d.nz = nz;
    d.X1 = nx + 2; d.Y1 = ny + 2; d.Z1 = nz + 2;
    d.X2 = nx + 4; d.Y2 = ny + 4; d.Z2 = nz + 4;
    d.X4 = nx + 8; d.Y4 = ny + 8; d.Z4 = nz + 8;
    d.S1 = (int64_t)d.X1 * d.Y1 * d.Z1;
    d.S2 = (int64_t)d.X2 * d.Y2 * d.Z2;
    d.S4 = (int64_t)d.X4 * d.Y4 * d.Z4;
    return d;
}

// i, j, k are 1-based interior coordinates (0 and n+1 are the first ghost layer).
MFLBM_HD int i1(const Dims& d, int i, int j, int k) { return i + d.X1 * (j + d.Y1 * k); }
MFLBM_HD int i2(const Dims& d, int i, int j, int k) { return (i + 1) + d.X2 * ((j + 1) + d.Y2 * (k + 1)); }
MFLBM_HD int i4(const Dims& d, int i, int j, int k) { return (i + 3) + d.X4 * ((j + 3) + d.Y4 * (k + 3)); }

// Node lists store packed coordinates: (i+4) | (j+4) << 10 | (k+4) << 21.
constexpr int PACK_BIAS = 4;
MFLBM_HD uint32_t pack_ijk(int i, int j, int k) {
    return (uint32_t)(i + PACK_BIAS) | ((uint32_t)(j + PACK_BIAS) << 10) | ((uint32_t)(k + PACK_BIAS) << 21);
}
MFLBM_HD void unpack_ijk(uint32_t p, int& i, int& j, int& k) {
    i = (int)(p & 1023u) - PACK_BIAS;
    j = (int)((p >> 10) & 2047u) - PACK_BIAS;
    k = (int)(p >> 21) - PACK_BIAS;
}

template <typename T>
struct Fields {
    Dims d;
    T* pdf;      // 38 * S1
    T* phi;      // S4
    T* cnx;      // S2
    T* cny;
    T* cnz;
    T* cnorm;    // S2
    T* curv;     // S1
    const T* snx;  // S4
    const T* sny;
    const T* snz;
    const uint8_t* walls;  // S2, 0/1
    const int8_t* wtype;   // S4, -1/0/1/2
};

template <typename T>
struct StepConst {
    T gamma, force_z, nui1, nui2, beta, phi_inlet, sa_inject, cos_theta, rho_in, rho_out, relaxation;
    T rho_ref;  // ShiftedStorage: density of the rest state the deviations are taken from
    T uin_avg;  // convective-outlet velocity
    int mrt;    // collision model 1-4 (mflbm_params::mrt)
};

struct NodeLists {
    const uint32_t* fluid;      // walls == 0 in [1,n]^3        : collide, curvature
    int n_fluid;
    const uint32_t* fluid_ext;  // walls == 0 in [-1,n+2]^3     : normals + wetting
    int n_fluid_ext;
    const uint32_t* solid_bnd;  // walls_type == 2 in [-2,n+3]^3: phi / normal extrapolation
    int n_solid_bnd;
};

// Rest state of the collision at rho = 1: (4/21, 13/126 x 6, 1/63 x 12) (MRT e = -11 rho,
// fourth-order moment 0).
MFLBM_HD double rest_weight(int q) { return q == 0 ? 4.0 / 21.0 : (q <= 6 ? 13.0 / 126.0 : 1.0 / 63.0); }

// Population storage of a precision.  FP64 keeps the reference's two colour populations
// (slot s: colour 1, slot s + 19: colour 2).  FP32 keeps colour 1 in slot s and, in slot
// s + 19, the total f_s - rho_ref * rest_weight(s): the deviation of the total population
// from the collision's rest state at the reference density rho_ref (StepConst; the mean
// density of the initial state).  Populations are O(0.1) while the flow lives in deviations of
// O(1e-6 .. 1e-3); a float holding the whole population resolves 7.5e-9 of it, a float
// holding the deviation resolves it to 6e-8 relative.  Density, momentum and the
// collision's moments are computed from the deviations, in which the rest state cancels
// exactly; colour 2 is f - colour 1 where it is needed.
template <typename T>
struct ShiftedStorage { static constexpr bool value = false; };
template <>
struct ShiftedStorage<float> { static constexpr bool value = true; };

constexpr int NB_WAVE = 40;  // words per wave of the short neighbour rows (36 bases, escape flag, padding)

// Fluid-compact population storage (see kernels.inl for the cell layout).
template <typename T>
struct Sparse {
    T* f;                           // 38 * ns: slot (s + 19 g) of every cell (ShiftedStorage), at at()
    int64_t ns;                     // cells (a multiple of CELL_BLOCK)
    int nf;                         // interior fluid nodes
    const int32_t* __restrict__ nb; // 18 * nf
    // The same links in half the bytes.  A wave of 64 nodes (t / 64) holds, per direction q,
    // the base cell of the fluid cells (class 0, cell < nf) and of the extra cells (class 1)
    // it links to: nbb[w * NB_WAVE + 2 (q - 1) + class]; nbb[w * NB_WAVE + 36] != 0 marks a
    // wave whose links of one class span 2^15 cells or more, which reads nb instead.  Each
    // node holds, per q, (cell - base) | class << 15 as 16 bits, directions 2p + 1 and 2p + 2
    // in the low and high half of nb16[p * nf + t].
    const uint32_t* __restrict__ nb16;  // 9 * nf
    const int32_t* __restrict__ nbb;    // NB_WAVE * ceil(nf / 64)
    const uint32_t* __restrict__ list;  // nf packed coordinates
    T* aux;                         // 4 * nf force records (cn_x, cn_y, cn_z, 0.5 gamma curv c_norm)
    // Position of slot s (0 .. 37) of cell c in f.  Cells come in blocks of CELL_BLOCK, each
    // block holding its 38 slots one after the other: a wave of consecutive cells reads
    // and writes 38 runs of 1 KiB (FP64) within one 38 KiB region rather than 38 separate
    // arrays (measured 5 % faster for the even collide than one array per slot).
    static constexpr int CELL_SHIFT = 7, CELL_BLOCK = 1 << CELL_SHIFT;
    MFLBM_HD int64_t at(int s, int64_t c) const {
        return ((c >> CELL_SHIFT) * 38 + s) * CELL_BLOCK + (c & (CELL_BLOCK - 1));
    }
};

// Inputs of the fused gradient-chain kernel (gradient.inl).  `tag` holds, for every node of
// the four-ghost-layer grid, (t << 2) | (walls_type & 3) with walls_type 0 -> 0, 1 -> 1,
// 2 -> 2, -1 -> 3, where t is the fluid-list index (force-record index) of an interior
// fluid node, the fluid-neighbour mask of a solid-boundary node (bit q-1 set when the D3Q19
// neighbour q is a fluid node, walls_type 0 or -1; static, like the geometry), 0 elsewhere.
template <typename T>
struct GradArgs {
    Dims d;
    T* phi;                           // S4
    const int32_t* __restrict__ tag;  // S4
    const T* __restrict__ snx;        // S4 solid normals, read at fluid-boundary nodes
    const T* __restrict__ sny;
    const T* __restrict__ snz;
    const T* __restrict__ snp;        // 4 * S4: (snx, sny, snz, 0) side by side (two-pass chain)
    T* aux;                           // 4 * nf force records
    T* cnx;                           // S2 dense normals and c_norm
    T* cny;
    T* cnz;
    T* cnorm;
    T* cnp;                           // 4 * S2: (cn_x, cn_y, cn_z, c_norm) side by side (two-pass chain)
    int shape, tiles_x, tiles_y, kchunk;
    // two-pass chain: boxes along z, and the order in which the workgroups walk the boxes
    // (0: x fastest, then y, then z; 1: x fastest, then z, then y)
    int tiles_z, box_order;
};

// Workgroup shapes of the fused gradient kernel: owned nodes in x and y, threads.
struct GradShape {
    int tx, ty, threads;
};
constexpr GradShape GRAD_SHAPES[] = {{16, 16, 256}, {16, 8, 256}, {8, 8, 128}, {16, 16, 384}, {32, 16, 512}};
constexpr int GRAD_NSHAPES = sizeof(GRAD_SHAPES) / sizeof(GRAD_SHAPES[0]);
constexpr int GRAD_SHAPE_DEFAULT = 0, GRAD_SHAPE_FP32 = 3;
// the two-pass chain on 3-D tiles (k_grad_normals, k_grad_force): the shape after the
// march shapes
constexpr int GRAD_SHAPE_TILES = GRAD_NSHAPES;
// owned nodes (x, y, z) of a workgroup of k_grad_normals (N) and k_grad_force (F), per precision
template <typename T>
struct GradTiles {
    static constexpr int N[3] = {16, 8, 4}, F[3] = {16, 8, 4};
};
template <>
struct GradTiles<float> {
    static constexpr int N[3] = {16, 8, 4}, F[3] = {16, 8, 4};
};

// Uniform deviate in [0, 1) of lattice node n under a seed, for the random initial phase
// distribution: the splitmix64 output function (Steele, Lea and Flood, OOPSLA 2014) of
// seed + (n + 1) * 0x9E3779B97F4A7C15, top 53 bits.  Counter-based: every node draws its own
// value, in any order, on the host or the device alike.
MFLBM_HD double node_uniform(uint64_t seed, uint64_t n) {
    uint64_t z = seed + (n + 1) * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return (double)(z >> 11) * 0x1p-53;
}

constexpr int TAG_FLUID = 0, TAG_SOLID = 1, TAG_SOLID_BND = 2, TAG_WET = 3;
MFLBM_HD bool tag_fluid(int code) { return code == TAG_FLUID || code == TAG_WET; }

struct BoundaryFlags {
    int kper, jper;
    int open_z;      // kper == 0 and no z walls
    int inlet_bc, outlet_bc;
};

}  // namespace mflbm
