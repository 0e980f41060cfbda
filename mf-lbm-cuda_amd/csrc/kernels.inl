// HIP kernels of the colour-gradient step, templated on the storage/arithmetic type T.
// Included by kernels_f64.hip (T = double, built with -ffp-contract=off so every
// operation rounds once, in source order, like the FP64 oracle) and kernels_f32.hip
// (T = float, contraction allowed).
//
// Reference correspondence (src/main_iteration_GPU.cu):
//   k_collide<ODD>     kernel_odd_color_GPU / kernel_even_color_GPU   (56-388 / 395-726)
//   k_periodic_pdf     perioidc_BC_{Z,Y,ZY}_{odd,even}                (1529-1689)
//   k_periodic_phi     periodic_phi_{Z,Y,ZY}                          (1691-1733)
//   k_open_bc<AFTER>   inlet/outlet_Zou_He_pressure_BC_{before,after}_odd_GPU (1085-1524),
//                      inlet_bounce_back_velocity_BC_* (1008-1077), outlet_convective_BC_* (1245-1357)
//   k_porous_plate     porous_plate_BC_{before,after}_odd (1744-1882)
//   k_phi_extrap       extrapolate_phi_toSolid                        (732-755)
//   k_normals          normalDirectionsOfInterfaces + alter_color_gradient_solid_surface_GPU
//                      (757-878), fused: the wetting correction only reads the node's own
//                      freshly computed normal
//   k_normal_extrap    extrapolateNormalToSolid                       (880-906)
//   k_curvature        CSF_Forces                                     (908-1003)
// The per-node arithmetic keeps the reference's expression order; that order is what
// makes the FP64 results bit-identical to the oracle.
#pragma once
#include <hip/hip_runtime.h>
#include <cmath>
#include <cstdint>
#include "lattice.h"
#include "zouhe.h"

namespace mflbm {
namespace dev {

// Workgroups are dispatched round-robin over the 8 XCDs of the MI355X.  Remap the
// hardware block index so that each XCD works on one contiguous range of the node list:
// neighbouring nodes (which share cache lines and stencil data) then meet in the same L2.
__device__ __forceinline__ int xcd_block() {
    const int nb = gridDim.x, b = blockIdx.x;
    const int q = nb >> 3, r = nb & 7, x = b & 7;
    return x * q + (x < r ? x : r) + (b >> 3);
}

// FP64 block reductions as wavefront-shuffle trees.  Each wave combines its 64 lanes with
// xor shuffles at distances 32, 16, ..., 1 (the two lanes of a pair compute the same
// commutative sum, so every lane ends with the wave's value), lane 0 parks it in LDS, and
// thread m combines the NT / 64 wave values of quantity m in wave order.  The pairing is
// fixed, so the result never depends on timing.  MAXI is the one quantity (if any) combined
// by fmax instead of +.  Thread m < NV receives quantity m in `out` and returns true; the
// others return false.  One call per kernel (its LDS is not guarded for reuse).
__device__ __forceinline__ double wave_tree(double v, bool is_max) {
#pragma unroll
    for (int m = 32; m > 0; m >>= 1) {
        const double o = __shfl_xor(v, m, 64);
        v = is_max ? fmax(v, o) : v + o;
    }
    return v;
}

template <int NV, int NT, int MAXI = -1>
__device__ __forceinline__ bool block_tree(double (&acc)[NV], double& out) {
    static_assert(NT % 64 == 0 && NV <= NT, "block of whole waves");
    __shared__ double wred[NV][NT / 64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int m = 0; m < NV; ++m) {
        const double v = wave_tree(acc[m], m == MAXI);
        if (lane == 0) wred[m][wave] = v;
    }
    __syncthreads();
    if ((int)threadIdx.x >= NV) return false;
    const int m = threadIdx.x;
    double s = wred[m][0];
    for (int w = 1; w < NT / 64; ++w) s = (m == MAXI) ? fmax(s, wred[m][w]) : s + wred[m][w];
    out = s;
    return true;
}

template <typename T>
struct W {
    static constexpr T w0 = T(1.0) / T(3.0);
    static constexpr T w1 = T(1.0) / T(18.0);
    static constexpr T w2 = T(1.0) / T(36.0);
    // rest state of the collision below: e = -11 rho and e2 = e2c1 rho = 0 give
    // f = rho (4/21, 13/126, 1/63) rather than the D3Q19 weights (whose e2 is 3 rho)
    static constexpr T r0 = T(4.0) / T(21.0);
    static constexpr T r1 = T(13.0) / T(126.0);
    static constexpr T r2 = T(1.0) / T(63.0);
};

template <typename T>
__device__ __forceinline__ T wq(int q) { return q == 0 ? W<T>::w0 : (q <= 6 ? W<T>::w1 : W<T>::w2); }

// MRT collision with Guo forcing of the total populations ft (src/main_iteration_GPU.cu:
// 164-345): moments, relaxation towards the equilibrium of density m_rho and velocity
// u = j + F/2, back-transform into f.  With ShiftedStorage ft are deviations from the rest
// state and m_rho the density deviation (see collide_node).
// The relaxation rates follow the collision model the reference selects with its `mrt`
// macro (src/main_iteration_GPU.cu:157-187): 1 bounce-back optimised, 2 original (the
// default), 3 SRT, 4 advection optimised; the model is uniform over a launch.
template <typename T>
__device__ __forceinline__ void mrt_collide(const T (&ft)[19], T m_rho, T fx, T fy, T fz, T omega, T (&f)[19],
                                            int mrt = 2) {
    T s_e = T(1.19), s_e2 = T(1.4), s_pi = T(1.4), s_q = T(1.2), s_t = T(1.98);
    if (mrt == 1) {
        s_e = s_e2 = s_pi = omega;
        s_q = s_t = T(8.0) * (T(2.0) - omega) / (T(8.0) - omega);
    } else if (mrt == 3) {
        s_e = s_e2 = s_pi = s_q = s_t = omega;
    } else if (mrt == 4) {
        s_e = s_e2 = s_pi = s_t = omega;
        s_q = (T(6.0) - T(3.0) * omega) / (T(3.0) - omega);
    }
    const T s_nu = omega;

    const T ux = ft[1] - ft[2] + ft[7] - ft[8] + ft[9] - ft[10] + ft[11] - ft[12] + ft[13] - ft[14] + T(0.5) * fx;
    const T uy = ft[3] - ft[4] + ft[7] + ft[8] - ft[9] - ft[10] + ft[15] - ft[16] + ft[17] - ft[18] + T(0.5) * fy;
    const T uz = ft[5] - ft[6] + ft[11] + ft[12] - ft[13] - ft[14] + ft[15] + ft[16] - ft[17] - ft[18] + T(0.5) * fz;
    const T u2 = ux * ux + uy * uy + uz * uz;

    T sum1 = ft[1] + ft[2] + ft[3] + ft[4] + ft[5] + ft[6];
    T sum2 = ft[7] + ft[8] + ft[9] + ft[10] + ft[11] + ft[12] + ft[13] + ft[14] + ft[15] + ft[16] + ft[17] + ft[18];
    T sum3 = ft[7] - ft[8] + ft[9] - ft[10] + ft[11] - ft[12] + ft[13] - ft[14];
    T sum4 = ft[7] + ft[8] - ft[9] - ft[10] + ft[15] - ft[16] + ft[17] - ft[18];
    T sum5 = ft[11] + ft[12] - ft[13] - ft[14] + ft[15] + ft[16] - ft[17] - ft[18];
    T sum6 = T(2.0) * (ft[1] + ft[2]) - ft[3] - ft[4] - ft[5] - ft[6];
    T sum7 = ft[7] + ft[8] + ft[9] + ft[10] + ft[11] + ft[12] + ft[13] + ft[14] - T(2.0) * (ft[15] + ft[16] + ft[17] + ft[18]);
    T sum8 = ft[3] + ft[4] - ft[5] - ft[6];
    T sum9 = ft[7] + ft[8] + ft[9] + ft[10] - ft[11] - ft[12] - ft[13] - ft[14];

    T m_e = T(-30.0) * ft[0] - T(11.0) * sum1 + T(8.0) * sum2;
    T m_e2 = T(12.0) * ft[0] - T(4.0) * sum1 + sum2;
    T m_jx = ft[1] - ft[2] + sum3;
    T m_qx = T(-4.0) * (ft[1] - ft[2]) + sum3;
    T m_jy = ft[3] - ft[4] + sum4;
    T m_qy = T(-4.0) * (ft[3] - ft[4]) + sum4;
    T m_jz = ft[5] - ft[6] + sum5;
    T m_qz = T(-4.0) * (ft[5] - ft[6]) + sum5;
    T m_3pxx = sum6 + sum7;
    T m_3pixx = T(-2.0) * sum6 + sum7;
    T m_pww = sum8 + sum9;
    T m_piww = T(-2.0) * sum8 + sum9;
    T m_pxy = ft[7] - ft[8] - ft[9] + ft[10];
    T m_pyz = ft[15] - ft[16] - ft[17] + ft[18];
    T m_pzx = ft[11] - ft[12] - ft[13] + ft[14];
    T m_tx = ft[7] - ft[8] + ft[9] - ft[10] - ft[11] + ft[12] - ft[13] + ft[14];
    T m_ty = -ft[7] - ft[8] + ft[9] + ft[10] + ft[15] - ft[16] + ft[17] - ft[18];
    T m_tz = ft[11] + ft[12] - ft[13] - ft[14] - ft[15] - ft[16] + ft[17] + ft[18];

    const T c23 = T(-0.666666666666666667), c13 = T(0.333333333333333333);
    const T e2c1 = T(0.0), e2c2 = T(-475.0) / T(63.0), oxx = T(0.0);
    m_e = m_e - s_e * (m_e - (T(-11.0) * m_rho + T(19.0) * u2)) + (T(38.0) - T(19.0) * s_e) * (fx * ux + fy * uy + fz * uz);
    m_e2 = m_e2 - s_e2 * (m_e2 - (e2c1 * m_rho + e2c2 * u2)) + (T(-11.0) + T(5.5) * s_e2) * (fx * ux + fy * uy + fz * uz);
    m_jx = m_jx + fx;
    m_qx = m_qx - s_q * (m_qx - (c23 * ux)) + (c23 + c13 * s_q) * fx;
    m_jy = m_jy + fy;
    m_qy = m_qy - s_q * (m_qy - (c23 * uy)) + (c23 + c13 * s_q) * fy;
    m_jz = m_jz + fz;
    m_qz = m_qz - s_q * (m_qz - (c23 * uz)) + (c23 + c13 * s_q) * fz;
    m_3pxx = m_3pxx - s_nu * (m_3pxx - (T(3.0) * ux * ux - u2)) + (T(2.0) - s_nu) * (T(2.0) * fx * ux - fy * uy - fz * uz);
    m_3pixx = m_3pixx - s_pi * (m_3pixx - oxx * (T(3.0) * ux * ux - u2)) + (T(1.0) - T(0.5) * s_pi) * (T(-2.0) * fx * ux + fy * uy + fz * uz);
    m_pww = m_pww - s_nu * (m_pww - (uy * uy - uz * uz)) + (T(2.0) - s_nu) * (fy * uy - fz * uz);
    m_piww = m_piww - s_pi * (m_piww - oxx * (uy * uy - uz * uz)) + (T(1.0) - T(0.5) * s_pi) * (-fy * uy + fz * uz);
    m_pxy = m_pxy - s_nu * (m_pxy - (ux * uy)) + (T(1.0) - T(0.5) * s_nu) * (fx * uy + fy * ux);
    m_pyz = m_pyz - s_nu * (m_pyz - (uy * uz)) + (T(1.0) - T(0.5) * s_nu) * (fy * uz + fz * uy);
    m_pzx = m_pzx - s_nu * (m_pzx - (ux * uz)) + (T(1.0) - T(0.5) * s_nu) * (fx * uz + fz * ux);
    m_tx = m_tx - s_t * (m_tx);
    m_ty = m_ty - s_t * (m_ty);
    m_tz = m_tz - s_t * (m_tz);

    const T c72 = T(1.0) / T(72.0);
    m_rho = (T(1.0) / T(19.0)) * m_rho;
    m_e = (T(1.0) / T(2394.0)) * m_e;
    m_e2 = (T(1.0) / T(252.0)) * m_e2;
    m_jx = T(0.1) * m_jx;  m_qx = T(0.025) * m_qx;
    m_jy = T(0.1) * m_jy;  m_qy = T(0.025) * m_qy;
    m_jz = T(0.1) * m_jz;  m_qz = T(0.025) * m_qz;
    m_3pxx = T(2.0) * c72 * m_3pxx;
    m_3pixx = c72 * m_3pixx;
    m_pww = T(6.0) * c72 * m_pww;
    m_piww = T(3.0) * c72 * m_piww;
    m_pxy = T(0.25) * m_pxy;  m_pyz = T(0.25) * m_pyz;  m_pzx = T(0.25) * m_pzx;
    m_tx = T(0.125) * m_tx;   m_ty = T(0.125) * m_ty;   m_tz = T(0.125) * m_tz;
    sum1 = m_rho - T(11.0) * m_e - T(4.0) * m_e2;
    sum2 = T(2.0) * m_3pxx - T(4.0) * m_3pixx;
    sum3 = m_pww - T(2.0) * m_piww;
    sum4 = m_rho + T(8.0) * m_e + m_e2;
    sum5 = m_jx + m_qx;
    sum6 = m_jy + m_qy;
    sum7 = m_jz + m_qz;
    sum8 = m_3pxx + m_3pixx;
    sum9 = m_pww + m_piww;
    f[0] = m_rho - T(30.0) * m_e + T(12.0) * m_e2;
    f[1] = sum1 + m_jx - T(4.0) * m_qx + sum2;
    f[2] = sum1 - m_jx + T(4.0) * m_qx + sum2;
    f[3] = sum1 + m_jy - T(4.0) * m_qy - T(0.5) * sum2 + sum3;
    f[4] = sum1 - m_jy + T(4.0) * m_qy - T(0.5) * sum2 + sum3;
    f[5] = sum1 + m_jz - T(4.0) * m_qz - T(0.5) * sum2 - sum3;
    f[6] = sum1 - m_jz + T(4.0) * m_qz - T(0.5) * sum2 - sum3;
    f[7] = sum4 + sum5 + sum6 + sum8 + sum9 + m_pxy + m_tx - m_ty;
    f[8] = sum4 - sum5 + sum6 + sum8 + sum9 - m_pxy - m_tx - m_ty;
    f[9] = sum4 + sum5 - sum6 + sum8 + sum9 - m_pxy + m_tx + m_ty;
    f[10] = sum4 - sum5 - sum6 + sum8 + sum9 + m_pxy - m_tx + m_ty;
    f[11] = sum4 + sum5 + sum7 + sum8 - sum9 + m_pzx - m_tx + m_tz;
    f[12] = sum4 - sum5 + sum7 + sum8 - sum9 - m_pzx + m_tx + m_tz;
    f[13] = sum4 + sum5 - sum7 + sum8 - sum9 - m_pzx - m_tx - m_tz;
    f[14] = sum4 - sum5 - sum7 + sum8 - sum9 + m_pzx + m_tx - m_tz;
    f[15] = sum4 + sum6 + sum7 - sum8 * T(2.0) + m_pyz + m_ty - m_tz;
    f[16] = sum4 - sum6 + sum7 - sum8 * T(2.0) - m_pyz - m_ty - m_tz;
    f[17] = sum4 + sum6 - sum7 - sum8 * T(2.0) - m_pyz + m_ty + m_tz;
    f[18] = sum4 - sum6 - sum7 - sum8 * T(2.0) + m_pyz - m_ty + m_tz;
}

// ---------------------------------------------------------------------------------------
// Collision + LKR recolouring of one node (src/main_iteration_GPU.cu:116-345).
// ---------------------------------------------------------------------------------------
// With ShiftedStorage (FP32) g2 / o2 hold the deviation of the total population from the
// rest state at rho_ref; the moments are then those of the deviation, in which only
// m_rho (rho_ref) and m_e (-11 rho_ref) of the rest state differ from zero, so the same
// relaxation applies with the density deviation in place of the density.
// ORIG: the reference's default model (mrt 2), whose constant rates the compiler folds; the
// other models read theirs from P.mrt in a kernel of their own, so that the default keeps
// its registers (the even collide needs 61 VGPRs in FP32 with constants, 118 without).
template <typename T, bool ORIG = true>
__device__ __forceinline__ T collide_node(const StepConst<T>& P, const T (&g1)[19], const T (&g2)[19], T cnx, T cny,
                                          T cnz, T ftmp, T (&o1)[19], T (&o2)[19]) {
    constexpr bool SH = ShiftedStorage<T>::value;
    T ft[19];
#pragma unroll
    for (int q = 0; q < 19; ++q) ft[q] = SH ? g2[q] : g1[q] + g2[q];
    T rho1 = g1[0], rho2 = g2[0];
#pragma unroll
    for (int q = 1; q < 19; ++q) { rho1 = rho1 + g1[q]; rho2 = rho2 + g2[q]; }
    T den, m_rho;
    if (SH) {
        m_rho = rho2;  // sum of the deviations
        den = P.rho_ref + m_rho;
        // Colour 1 is stored whole, so where fluid 2 is absent colour 2 = total - colour 1
        // is a cancellation that leaves the float rounding of colour 1 as a spurious trace
        // of fluid 2.  Where fluid 2 is absent, every colour-1 slot was rounded once from
        // the total it is stored with (g1_q = fl(D_q + rho_ref r_q), by the recolouring
        // below at rho2 = 0, by the initialisation, by the boundary accessors), so the
        // exact residue rho_ref + sum D_q - sum g1_q is at most 2^-24 sum |g1_q| plus the
        // rounding of the rest state, under 2^-23 sum |g1_q|.  Summed exactly (in FP64),
        // a residue within 2^-22 sum |g1_q| is that rounding and is read as no fluid 2,
        // which keeps such regions at phi = 1 exactly; anything larger, of either sign,
        // is a real trace of fluid 2 and is kept, as the FP64 build keeps it.  (Where fluid
        // 1 is absent colour 1 is exactly zero and nothing needs reading.)
        double r1d = 0.0, a1d = 0.0, td = (double)P.rho_ref;
#pragma unroll
        for (int q = 0; q < 19; ++q) {
            r1d += (double)g1[q];
            a1d += fabs((double)g1[q]);
            td += (double)g2[q];
        }
        if (fabs(td - r1d) <= 0x1p-22 * a1d) rho1 = den;
        rho2 = den - rho1;
    } else {
        den = rho1 + rho2;
        m_rho = den;
    }
    const T phi = (rho1 - rho2) / (SH ? den : rho1 + rho2);

    // ftmp = 0.5 * gamma * curv * c_norm, evaluated in that order by k_curvature
    T tmp = ftmp;
    const T fx = tmp * cnx, fy = tmp * cny, fz = tmp * cnz + P.force_z;

    const T omega = T(1.0) / (T(6.0) / ((T(1.0) + phi) * P.nui1 + (T(1.0) - phi) * P.nui2) + T(0.5));
    T f[19];
    mrt_collide<T>(ft, m_rho, fx, fy, fz, omega, f, ORIG ? 2 : P.mrt);

    const T tmp1 = rho1 / den;
    o1[0] = tmp1 * (SH ? f[0] + P.rho_ref * T(rest_weight(0)) : f[0]);
    o2[0] = SH ? f[0] : f[0] * (T(1.0) - tmp1);
    tmp = rho1 * rho2 * P.beta / den;
    const T rk2 = T(1.0) / sqrt(T(2.0)) / T(36.0);
    const T w1 = T(1.0) / T(18.0);
    const T pr[19] = {T(0), cnx, -cnx, cny, -cny, cnz, -cnz,
                      cnx + cny, -cnx + cny, cnx - cny, -cnx - cny,
                      cnx + cnz, -cnx + cnz, cnx - cnz, -cnx - cnz,
                      cny + cnz, -cny + cnz, cny - cnz, -cny - cnz};
#pragma unroll
    for (int q = 1; q < 19; ++q) {
        const T w = q <= 6 ? w1 : rk2;
        o1[q] = tmp1 * (SH ? f[q] + P.rho_ref * T(rest_weight(q)) : f[q]) + w * tmp * (pr[q]);
        o2[q] = SH ? f[q] : f[q] - o1[q];
    }
    return phi;
}


// ---------------------------------------------------------------------------------------
// Fluid-compact population storage.
// Cells hold, per slot, the interior fluid nodes first (index = position in the fluid list)
// followed by the slot's extra cells: every non-fluid (node, slot) pair a step ever reads or
// writes (bounce-back buffers in solid neighbours, ghost and periodic image cells, cells read
// by the open boundaries); S.at(slot, cell) places them.  nb[(q-1)*nf + t] is the cell of
// (x_t + e_q, slot opp(q)): an odd step pushes direction q into that cell and pulls
// direction opp(q) out of it (the same cell, hence in place).  Fluid neighbours have the
// same cell in every slot (their list position).
// ---------------------------------------------------------------------------------------

// AA-pattern collide/stream of one interior fluid node (src/main_iteration_GPU.cu:56-388
// odd, 395-726 even).  Odd: pull q from (x - e_q).slot q, push q to (x + e_q).slot opp(q);
// even: read q from x.slot opp(q), write q to x.slot q.  Each cell is read and written by
// exactly one node per sweep, so the update is in place.
// Links of node t (cells of x + e_q, q = 1..18, into nb[1..18]) from the short neighbour rows
// (Sparse<T>::nb16 / nbb): the wave's bases, one address for the whole wave (scalar loads),
// and the node's nine pairs of 16-bit links; a flagged wave reads the 32-bit rows nb32
// instead.  38.5 bytes per node against 72.  t / 64 must be the wave's index, i.e. nodes
// are mapped to lanes in order.  The loads are issued, and held in front of what follows,
// before the links are decoded.
template <typename I>
__device__ __forceinline__ void odd_links(const int32_t* __restrict__ nbb, const uint32_t* __restrict__ nb16,
                                          const int32_t* __restrict__ nb32, int64_t nf, int64_t t, I (&nb)[19]) {
    const int w = __builtin_amdgcn_readfirstlane((int)(t >> 6));
    const int4* wb = reinterpret_cast<const int4*>(nbb + (int64_t)w * NB_WAVE);
    int4 bq[NB_WAVE / 4];
#pragma unroll
    for (int m = 0; m < NB_WAVE / 4; ++m) bq[m] = wb[m];
    uint32_t pr[9];
#pragma unroll
    for (int p = 0; p < 9; ++p) pr[p] = __builtin_nontemporal_load(nb16 + (int64_t)p * nf + t);
    __builtin_amdgcn_sched_barrier(0);
    if (bq[9].x) {  // a wave whose links span too many cells for 15 bits
#pragma unroll
        for (int q = 1; q < 19; ++q) nb[q] = (I)__builtin_nontemporal_load(nb32 + (int64_t)(q - 1) * nf + t);
    } else {
        auto word = [&](int k) {
            const int4& v = bq[k >> 2];
            return (k & 3) == 0 ? v.x : (k & 3) == 1 ? v.y : (k & 3) == 2 ? v.z : v.w;
        };
#pragma unroll
        for (int q = 1; q < 19; ++q) {
            const uint32_t u = ((q - 1) & 1) ? pr[(q - 1) >> 1] >> 16 : pr[(q - 1) >> 1] & 0xFFFFu;
            const int k = 2 * (q - 1);
            nb[q] = (I)(((u >> 15) ? word(k + 1) : word(k)) + (int)(u & 0x7FFFu));
        }
    }
}

#ifndef MFLBM_COLLIDE_THREADS
#define MFLBM_COLLIDE_THREADS 256  // threads per collide workgroup (tuning: 128, 256, 512)
#endif
template <typename T, bool ODD, bool ORIG = true>
__global__ __launch_bounds__(MFLBM_COLLIDE_THREADS) void k_collide(Sparse<T> S, T* __restrict__ phi, Dims d,
                                                                   StepConst<T> P) {
    const int t = xcd_block() * MFLBM_COLLIDE_THREADS + threadIdx.x;
    if (t >= S.nf) return;
    // The neighbour row, force record and phi of the odd step are touched once; nontemporal
    // accesses leave the cache to the shared lines of the scattered cells (6 % faster; in
    // the even kernel they cost 2-5 %).  The odd step loads the node's own words (force
    // record, packed coordinates) with its neighbour row, held there by a scheduling
    // barrier: sunk to their first use they were issued after the population loads had
    // returned, a third round trip to memory (0.6 % on the odd launch; the same order in
    // the even kernel measured 0.8 % slower and is not used there).
    const T* a = S.aux + 4 * (int64_t)t;
    T rec[4];
    uint32_t packed = 0;
    T g1[19], g2[19], o1[19], o2[19];
    int nb[19];
    if (ODD) {
#pragma unroll
        for (int m = 0; m < 4; ++m) rec[m] = __builtin_nontemporal_load(a + m);
        packed = S.list[t];
        nb[0] = t;
        odd_links(S.nbb, S.nb16, S.nb, S.nf, t, nb);
    }
#pragma unroll
    for (int q = 0; q < 19; ++q) {
        const int64_t src = ODD ? (int64_t)nb[OPP[q]] : (int64_t)t;
        const int slot = ODD ? q : OPP[q];
        // even: every cell is the node's own and streams through once, so nontemporal loads
        // and stores (18-24 % faster); odd: neighbouring nodes share the cache lines of their
        // scattered cells, which nontemporal accesses would fetch again (40-70 % slower)
        if (ODD) {
            g1[q] = S.f[S.at(slot, src)];
            g2[q] = S.f[S.at(slot + 19, src)];
        } else {
            g1[q] = __builtin_nontemporal_load(S.f + S.at(slot, src));
            g2[q] = __builtin_nontemporal_load(S.f + S.at(slot + 19, src));
        }
    }
    if (!ODD) {
#pragma unroll
        for (int m = 0; m < 4; ++m) rec[m] = a[m];
        packed = S.list[t];
    }
    const T ph = collide_node<T, ORIG>(P, g1, g2, rec[0], rec[1], rec[2], rec[3], o1, o2);
    int i, j, k;
    unpack_ijk(packed, i, j, k);
    if (ODD) __builtin_nontemporal_store(ph, phi + i4(d, i, j, k));
    else phi[i4(d, i, j, k)] = ph;
#pragma unroll
    for (int q = 0; q < 19; ++q) {
        const int64_t dst = ODD ? (int64_t)nb[q] : (int64_t)t;
        const int slot = ODD ? OPP[q] : q;
        if (ODD) {
            S.f[S.at(slot, dst)] = o1[q];
            S.f[S.at(slot + 19, dst)] = o2[q];
        } else {
            __builtin_nontemporal_store(o1[q], S.f + S.at(slot, dst));
            __builtin_nontemporal_store(o2[q], S.f + S.at(slot + 19, dst));
        }
    }
}

// Periodic population copies, precomputed as (slot, destination cell, source cell) in the
// reference's kernel order with repeated destinations resolved to the last writer
// (perioidc_BC_{Z,Y,ZY}_{odd,even}, src/main_iteration_GPU.cu:1529-1689).
template <typename T>
__global__ __launch_bounds__(256) void k_periodic_pdf(Sparse<T> S, const int32_t* __restrict__ pairs, int n) {
    const int t = blockIdx.x * 256 + threadIdx.x;
    if (t >= n) return;
    const int slot = pairs[3 * t], dst = pairs[3 * t + 1], src = pairs[3 * t + 2];
    S.f[S.at(slot, dst)] = S.f[S.at(slot, src)];
    S.f[S.at(slot + 19, dst)] = S.f[S.at(slot + 19, src)];
}

// Periodic ghost layers of phi (periodic_phi_{Z,Y,ZY}, src/main_iteration_GPU.cu:1691-1733);
// the three groups read interior cells and write disjoint ghost cells.
template <typename T>
__global__ __launch_bounds__(256) void k_periodic_phi(T* __restrict__ phi, Dims d, int kper, int jper) {
    const int nx = d.nx, ny = d.ny, nz = d.nz;
    int t = blockIdx.x * 256 + threadIdx.x;
    auto P = [&](int a, int b, int c) -> T& { return phi[i4(d, a, b, c)]; };
    const int nzf = kper ? nx * ny : 0, nyf = jper ? nx * nz : 0, ne = (kper && jper) ? nx : 0;
    if (t < nzf) {
        const int i = t % nx + 1, j = t / nx + 1;
#pragma unroll
        for (int kk = 1; kk <= 4; ++kk) {
            P(i, j, kk + nz) = P(i, j, kk);
            P(i, j, kk - 4) = P(i, j, nz + kk - 4);
        }
        return;
    }
    t -= nzf;
    if (t < nyf) {
        const int i = t % nx + 1, k = t / nx + 1;
#pragma unroll
        for (int jj = 1; jj <= 4; ++jj) {
            P(i, jj + ny, k) = P(i, jj, k);
            P(i, jj - 4, k) = P(i, ny + jj - 4, k);
        }
        return;
    }
    t -= nyf;
    if (t < ne) {
        const int i = t + 1;
        for (int kk = 1; kk <= 4; ++kk)
            for (int jj = 1; jj <= 4; ++jj) {
                P(i, jj - 4, kk - 4) = P(i, ny + jj - 4, nz + kk - 4);
                P(i, jj + ny, kk - 4) = P(i, jj, nz + kk - 4);
                P(i, jj + ny, kk + nz) = P(i, jj, kk);
                P(i, jj - 4, kk + nz) = P(i, ny + jj - 4, kk);
            }
    }
}

// One colour population of a cell in shifted storage (colour 1 and the total's deviation
// from the rest state), read and written as the plain colour value; writing colour 1 keeps
// colour 2 unchanged.
template <typename T>
struct ColourRef {
    T* g1;
    T* dev;
    T rest;
    bool colour2;
    __device__ operator T() const { return colour2 ? (*dev + rest) - *g1 : *g1; }
    __device__ ColourRef& operator=(T v) {
        if (colour2) {
            *dev = (*g1 + v) - rest;
        } else {
            const T g2 = (*dev + rest) - *g1;
            *g1 = v;
            *dev = (v + g2) - rest;
        }
        return *this;
    }
};

// Runs fn(f) with the population accessor of the storage: f(a, b, c, e, g) is a plain
// reference for the two-colour storage and a ColourRef for the deviation storage; cell_of
// maps (a, b, c, e) to the cell index.
template <typename T, class CellOf, class Fn>
__device__ __forceinline__ void with_cells(const Sparse<T>& S, T rho_ref, CellOf&& cell_of, Fn&& fn) {
    if constexpr (ShiftedStorage<T>::value) {
        auto f = [&](int a, int b, int c, int e, int g) -> ColourRef<T> {
            const int64_t cell = cell_of(a, b, c, e);
            return ColourRef<T>{S.f + S.at(e, cell), S.f + S.at(e + 19, cell), rho_ref * (T)rest_weight(e), g == 1};
        };
        fn(f);
    } else {
        auto f = [&](int a, int b, int c, int e, int g) -> T& { return S.f[S.at(e + 19 * g, cell_of(a, b, c, e))]; };
        fn(f);
    }
}

// Open inlet/outlet, one thread per (i, j): in_bc / out_bc 2 = Zou-He, 1 = velocity inlet /
// convective outlet.  zmap[plane][slot][b * (nx+2) + a] gives the cell of (a, b, k) for the
// six planes k = 0, 1, 2, nz-1, nz, nz+1.  w_in: the inlet velocity plane (X1 * Y1);
// conv: the convective-outlet state, 38 planes of populations [g][e] then phi.
template <typename T, bool AFTER_ODD>
__global__ __launch_bounds__(256) void k_open_bc(Sparse<T> S, Fields<T> F, StepConst<T> P, const int32_t* __restrict__ zmap,
                                                 int in_bc, int out_bc, const T* __restrict__ w_in, T* __restrict__ conv,
                                                 int split) {
    const Dims& d = F.d;
    int t = blockIdx.x * 256 + threadIdx.x;
    // split: the inlet columns and the outlet columns on threads of their own (the two
    // planes share no cell); otherwise one thread does both ends of its column
    if (t >= (split ? 2 : 1) * d.nx * d.ny) return;
    if (split) {
        if (t < d.nx * d.ny) out_bc = 0;
        else { in_bc = 0; t -= d.nx * d.ny; }
    }
    const int i = t % d.nx + 1, j = t / d.nx + 1, nz = d.nz;
    const int64_t P2 = (int64_t)d.X1 * d.Y1, plane = 19 * P2;
    auto cell_of = [&](int a, int b, int c, int e) -> int64_t {
        const int pl = c <= 2 ? c : 3 + (c - (nz - 1));
        return zmap[pl * plane + (int64_t)e * P2 + b * d.X1 + a];
    };
    auto wall = [&](int a, int b, int c) -> int { return F.walls[i2(d, a, b, c)]; };
    auto ph = [&](int a, int b, int c) -> T& { return F.phi[i4(d, a, b, c)]; };
    const int64_t col = (int64_t)j * d.X1 + i;
    with_cells<T>(S, P.rho_ref, cell_of, [&](auto&& f) {
        if (in_bc == 1) velocity_inlet_column<T, AFTER_ODD>(P, i, j, w_in[col], f, wall, ph);
        if (in_bc == 2 || out_bc == 2) zou_he_column<T, AFTER_ODD>(P, i, j, nz, in_bc == 2, out_bc == 2, f, wall, ph);
        if (out_bc == 1) {
            auto cs = [&](int e, int g) -> T& { return conv[(int64_t)(e + 19 * g) * P2 + col]; };
            auto cp = [&]() -> T& { return conv[38 * P2 + col]; };
            convective_outlet_column<T, AFTER_ODD>(P, i, j, nz, f, wall, ph, cs, cp);
        }
    });
}

// Porous plate at k = zp, one thread per (i, j); pmap as zmap for the planes zp-1, zp, zp+1.
template <typename T, bool AFTER_ODD>
__global__ __launch_bounds__(256) void k_porous_plate(Sparse<T> S, Dims d, StepConst<T> P, const int32_t* __restrict__ pmap,
                                                      int zp, int blocked) {
    const int t = blockIdx.x * 256 + threadIdx.x;
    if (t >= d.nx * d.ny) return;
    const int i = t % d.nx + 1, j = t / d.nx + 1;
    const int64_t P2 = (int64_t)d.X1 * d.Y1, plane = 19 * P2;
    auto cell_of = [&](int a, int b, int c, int e) -> int64_t {
        return pmap[(c - (zp - 1)) * plane + (int64_t)e * P2 + b * d.X1 + a];
    };
    with_cells<T>(S, P.rho_ref, cell_of, [&](auto&& f) { porous_plate_column<AFTER_ODD>(i, j, zp, blocked, f); });
}

// Convective-outlet state of a new run: the populations and phi of the outlet plane k = nz
// of the initial state (src/Init_multiphase.cpp:445-494), in plain colour values.
template <typename T>
__global__ __launch_bounds__(256) void k_conv_init(Fields<T> F, T rho_ref, T* __restrict__ conv) {
    const Dims& d = F.d;
    const int64_t P2 = (int64_t)d.X1 * d.Y1;
    const int64_t col = blockIdx.x * 256 + threadIdx.x;
    if (col >= P2) return;
    const int64_t n = col + P2 * d.nz;  // (a, b, nz) in the s1 grid
    for (int e = 0; e < 19; ++e) {
        T g1 = F.pdf[n + d.S1 * e], g2 = F.pdf[n + d.S1 * (e + 19)];
        if (ShiftedStorage<T>::value) g2 = (g2 + rho_ref * (T)rest_weight(e)) - g1;
        conv[(int64_t)e * P2 + col] = g1;
        conv[(int64_t)(e + 19) * P2 + col] = g2;
    }
    const int a = (int)(col % d.X1), b = (int)(col / d.X1);
    conv[38 * P2 + col] = F.phi[i4(d, a, b, d.nz)];
}

// Dense reference layout <-> compact storage.  Fluid cells: t < nf; extra cells: the
// ext_node list of each slot (offsets ext_off[0..19]).
template <typename T, bool TO_DENSE>
__global__ __launch_bounds__(256) void k_fluid_cells(Sparse<T> S, T* __restrict__ dense, Dims d) {
    const int t = blockIdx.x * 256 + threadIdx.x;
    if (t >= S.nf) return;
    int i, j, k;
    unpack_ijk(S.list[t], i, j, k);
    const int64_t c = i1(d, i, j, k);
    for (int s = 0; s < 38; ++s) {
        if (TO_DENSE) dense[s * d.S1 + c] = S.f[S.at(s, t)];
        else S.f[S.at(s, t)] = dense[s * d.S1 + c];
    }
}

template <typename T, bool TO_DENSE>
__global__ __launch_bounds__(256) void k_extra_cells(Sparse<T> S, T* __restrict__ dense, Dims d,
                                                     const int32_t* __restrict__ ext_node,
                                                     const int32_t* __restrict__ ext_off) {
    const int t = blockIdx.x * 256 + threadIdx.x;
    if (t >= ext_off[19]) return;
    int s = 0;
    while (t >= ext_off[s + 1]) ++s;
    const int e = t - ext_off[s];
    const int64_t c = ext_node[t];
    for (int g = 0; g < 2; ++g) {
        const int sl = s + 19 * g;
        if (TO_DENSE) dense[sl * d.S1 + c] = S.f[S.at(sl, S.nf + e)];
        else S.f[S.at(sl, S.nf + e)] = dense[sl * d.S1 + c];
    }
}

// ---------------------------------------------------------------------------------------
// Colour-gradient chain
// ---------------------------------------------------------------------------------------
// ISO4 central-difference stencil along one axis, summed in the reference order
// (src/main_iteration_GPU.cu:765-791, 914-996).  `a(dx, dy, dz)` returns the value at the
// neighbour offset; the dense-array and the LDS-tile callers share this one body so they
// round identically.
template <typename T, int AX, class A>
__device__ __forceinline__ T iso4_at(A&& a) {
    // offsets of the 4 "plus" terms of the second group, per axis
    constexpr int G2[3][4][3] = {{{1, 1, 0}, {1, -1, 0}, {1, 0, 1}, {1, 0, -1}},
                                 {{1, 1, 0}, {-1, 1, 0}, {0, 1, 1}, {0, 1, -1}},
                                 {{1, 0, 1}, {-1, 0, 1}, {0, 1, 1}, {0, -1, 1}}};
    constexpr int ax = AX == 0, ay = AX == 1, az = AX == 2;
    const T d1 = a(ax, ay, az) - a(-ax, -ay, -az);
    T acc = T(0);
#pragma unroll
    for (int m = 0; m < 4; ++m) {
        const int ox = G2[AX][m][0], oy = G2[AX][m][1], oz = G2[AX][m][2];
        acc = (m == 0) ? a(ox, oy, oz) : acc + a(ox, oy, oz);
        acc = acc - a(-ox, -oy, -oz);
    }
    return (T(1.0) / T(6.0)) * d1 + (T(1.0) / T(12.0)) * acc;
}

template <typename T, int AX>
__device__ __forceinline__ T iso4(const T* __restrict__ a, int c, int sx, int sy, int sz) {
    return iso4_at<T, AX>([&](int dx, int dy, int dz) -> T { return a[c + dx * sx + dy * sy + dz * sz]; });
}

// Interface normal of a fluid node from its ISO4 colour gradient (cx, cy, cz), with the
// geometrical wetting correction at fluid-boundary nodes (wetting = walls_type -1) against
// the solid normal (nwx, nwy, nwz): src/main_iteration_GPU.cu:793-807 and 809-878.
// Returns c_norm; (vx, vy, vz) receive the unit normal (zero where |grad| < 1e-6).
template <typename T, class SolidNormal>
__device__ __forceinline__ T interface_normal(T cx, T cy, T cz, bool wetting, T ct, SolidNormal&& solid, T& vx, T& vy,
                                              T& vz) {
    const T cn = sqrt(cx * cx + cy * cy + cz * cz);
    if (cn < T(1e-6)) {
        vx = T(0); vy = T(0); vz = T(0);
        return T(0);
    }
    T vx0 = cx / cn, vy0 = cy / cn, vz0 = cz / cn;
    if (wetting && cn > T(1e-6)) {
        const T lambda = T(0.5), leps = T(1e-6);
        T nwx, nwy, nwz;
        solid(nwx, nwy, nwz);
        T vx1 = vx0 - lambda * (vx0 + nwx);
        T vy1 = vy0 - lambda * (vy0 + nwy);
        T vz1 = vz0 - lambda * (vz0 + nwz);
        T e0 = (nwx * vx0 + nwy * vy0 + nwz * vz0) - ct;
        if ((fabs(vx0 + nwx) + fabs(vy0 + nwy) + fabs(vz0 + nwz) > leps ||
             fabs(vx0 - nwx) + fabs(vy0 - nwy) + fabs(vz0 - nwz) > leps) && e0 > leps) {
            T e1 = (nwx * vx1 + nwy * vy1 + nwz * vz1) - sqrt(vx1 * vx1 + vy1 * vy1 + vz1 * vz1) * ct;
            T tt = T(1.0) / (e1 - e0);
            T vx2 = tt * (vx0 * e1 - vx1 * e0);
            T vy2 = tt * (vy0 * e1 - vy1 * e0);
            T vz2 = tt * (vz0 * e1 - vz1 * e0);
            T e2 = (nwx * vx2 + nwy * vy2 + nwz * vz2) - sqrt(vx2 * vx2 + vy2 * vy2 + vz2 * vz2) * ct;
            if (e2 > leps) {
                for (int it = 2; it <= 4; ++it) {
                    vx0 = vx1; vy0 = vy1; vz0 = vz1;
                    vx1 = vx2; vy1 = vy2; vz1 = vz2;
                    e0 = (nwx * vx0 + nwy * vy0 + nwz * vz0) - sqrt(vx0 * vx0 + vy0 * vy0 + vz0 * vz0) * ct;
                    e1 = (nwx * vx1 + nwy * vy1 + nwz * vz1) - sqrt(vx1 * vx1 + vy1 * vy1 + vz1 * vz1) * ct;
                    tt = T(1.0) / (e1 - e0);
                    if (isinf(tt)) break;
                    vx2 = tt * (vx0 * e1 - vx1 * e0);
                    vy2 = tt * (vy0 * e1 - vy1 * e0);
                    vz2 = tt * (vz0 * e1 - vz1 * e0);
                    e2 = (nwx * vx2 + nwy * vy2 + nwz * vz2) - sqrt(vx2 * vx2 + vy2 * vy2 + vz2 * vz2) * ct;
                }
            }
            tt = T(1.0) / (T(1e-30) + sqrt(vx2 * vx2 + vy2 * vy2 + vz2 * vz2));
            vx0 = vx2 * tt; vy0 = vy2 * tt; vz0 = vz2 * tt;
        }
    }
    vx = vx0; vy = vy0; vz = vz0;
    return cn;
}

// Curvature of the interface from the ISO4 derivatives of the unit normal field
// (src/main_iteration_GPU.cu:914-1001).  `nx/ny/nz(dx, dy, dz)` read the normal
// components at a neighbour offset.
template <typename T, class NX, class NY, class NZ>
__device__ __forceinline__ T curvature_of(NX&& nx, NY&& ny, NZ&& nz) {
    const T kxx = iso4_at<T, 0>(nx);
    const T kyy = iso4_at<T, 1>(ny);
    const T kzz = iso4_at<T, 2>(nz);
    const T kxy = iso4_at<T, 1>(nx);
    const T kxz = iso4_at<T, 2>(nx);
    const T kyx = iso4_at<T, 0>(ny);
    const T kyz = iso4_at<T, 2>(ny);
    const T kzx = iso4_at<T, 0>(nz);
    const T kzy = iso4_at<T, 1>(nz);
    const T cx = nx(0, 0, 0), cy = ny(0, 0, 0), cz = nz(0, 0, 0);
    return (cx * cx - T(1.0)) * kxx + (cy * cy - T(1.0)) * kyy + (cz * cz - T(1.0)) * kzz + cx * cy * (kxy + kyx) +
           cx * cz * (kxz + kzx) + cy * cz * (kzy + kyz);
}

// K3: phi at solid-boundary nodes = D3Q19-weighted mean over fluid neighbours, [-2, n+3].
template <typename T>
__global__ __launch_bounds__(256) void k_phi_extrap(Fields<T> F, const uint32_t* __restrict__ list, int n) {
    const int t = xcd_block() * 256 + threadIdx.x;
    if (t >= n) return;
    int i, j, k;
    unpack_ijk(list[t], i, j, k);
    const Dims& d = F.d;
    const int c = i4(d, i, j, k);
    T ps = T(0), ws = T(0);
#pragma unroll
    for (int q = 1; q < 19; ++q) {
        const int nb = c + EX[q] + d.X4 * (EY[q] + d.Y4 * EZ[q]);
        if (F.wtype[nb] <= 0) {
            ps += F.phi[nb] * wq<T>(q);
            ws += wq<T>(q);
        }
    }
    F.phi[c] = ps / ws;
}

// K4 + K5: interface normals from phi (ISO4), then the geometrical wetting correction at
// fluid-boundary nodes, over fluid nodes in [-1, n+2].  Solid nodes keep cn = c_norm = 0
// from initialisation (the reference rewrites those zeros every step; K6 then overwrites
// the solid-boundary normals in [0, n+1] from fluid neighbours only).
template <typename T>
__global__ __launch_bounds__(256) void k_normals(Fields<T> F, StepConst<T> P, const uint32_t* __restrict__ list, int n) {
    const int t = xcd_block() * 256 + threadIdx.x;
    if (t >= n) return;
    int i, j, k;
    unpack_ijk(list[t], i, j, k);
    const Dims& d = F.d;
    const int c4 = i4(d, i, j, k), c2 = i2(d, i, j, k);
    const int sx = 1, sy = d.X4, sz = d.X4 * d.Y4;
    const T cx = iso4<T, 0>(F.phi, c4, sx, sy, sz);
    const T cy = iso4<T, 1>(F.phi, c4, sx, sy, sz);
    const T cz = iso4<T, 2>(F.phi, c4, sx, sy, sz);
    T vx, vy, vz;
    F.cnorm[c2] = interface_normal<T>(cx, cy, cz, F.wtype[c4] == -1, P.cos_theta,
                                      [&](T& nwx, T& nwy, T& nwz) { nwx = F.snx[c4]; nwy = F.sny[c4]; nwz = F.snz[c4]; },
                                      vx, vy, vz);
    F.cnx[c2] = vx; F.cny[c2] = vy; F.cnz[c2] = vz;
}

// K6: normals at solid-boundary nodes in [0, n+1] from fluid neighbours.
template <typename T>
__global__ __launch_bounds__(256) void k_normal_extrap(Fields<T> F, const uint32_t* __restrict__ list, int n) {
    const int t = xcd_block() * 256 + threadIdx.x;
    if (t >= n) return;
    int i, j, k;
    unpack_ijk(list[t], i, j, k);
    const Dims& d = F.d;
    if (i < 0 || j < 0 || k < 0 || i > d.nx + 1 || j > d.ny + 1 || k > d.nz + 1) return;
    const int c4 = i4(d, i, j, k), c2 = i2(d, i, j, k);
    T sx = T(0), sy = T(0), sz = T(0), ws = T(0);
#pragma unroll
    for (int q = 1; q < 19; ++q) {
        if (F.wtype[c4 + EX[q] + d.X4 * (EY[q] + d.Y4 * EZ[q])] <= 0) {
            const int nb = c2 + EX[q] + d.X2 * (EY[q] + d.Y2 * EZ[q]);
            sx += F.cnx[nb] * wq<T>(q);
            sy += F.cny[nb] * wq<T>(q);
            sz += F.cnz[nb] * wq<T>(q);
            ws += wq<T>(q);
        }
    }
    F.cnx[c2] = sx / ws; F.cny[c2] = sy / ws; F.cnz[c2] = sz / ws;
}

// K7: curvature from the ISO4 derivatives of the normals, at interior fluid nodes.  Also
// writes the node's record for the next collide, in fluid-list order:
// (cn_x, cn_y, cn_z, 0.5 * gamma * curv * c_norm).
template <typename T>
__device__ __forceinline__ T curvature_at(const Fields<T>& F, int c2) {
    const Dims& d = F.d;
    const int sy = d.X2, sz = d.X2 * d.Y2;
    auto at = [&](const T* a) {
        return [=](int dx, int dy, int dz) -> T { return a[c2 + dx + dy * sy + dz * sz]; };
    };
    return curvature_of<T>(at(F.cnx), at(F.cny), at(F.cnz));
}

template <typename T>
__global__ __launch_bounds__(256) void k_curvature(Fields<T> F, StepConst<T> P, const uint32_t* __restrict__ list, int n,
                                                   T* __restrict__ aux) {
    const int t = xcd_block() * 256 + threadIdx.x;
    if (t >= n) return;
    int i, j, k;
    unpack_ijk(list[t], i, j, k);
    const Dims& d = F.d;
    const int c2 = i2(d, i, j, k);
    const T curv = curvature_at(F, c2);
    F.curv[i1(d, i, j, k)] = curv;
    aux[4 * (int64_t)t + 0] = F.cnx[c2];
    aux[4 * (int64_t)t + 1] = F.cny[c2];
    aux[4 * (int64_t)t + 2] = F.cnz[c2];
    aux[4 * (int64_t)t + 3] = T(0.5) * P.gamma * curv * F.cnorm[c2];
}

// Rebuilds the force records from the stored normals and curvature (after an upload).
template <typename T>
__global__ __launch_bounds__(256) void k_aux_refresh(Fields<T> F, StepConst<T> P, const uint32_t* __restrict__ list,
                                                     int n, T* __restrict__ aux) {
    const int t = blockIdx.x * 256 + threadIdx.x;
    if (t >= n) return;
    int i, j, k;
    unpack_ijk(list[t], i, j, k);
    const Dims& d = F.d;
    const int c2 = i2(d, i, j, k);
    aux[4 * (int64_t)t + 0] = F.cnx[c2];
    aux[4 * (int64_t)t + 1] = F.cny[c2];
    aux[4 * (int64_t)t + 2] = F.cnz[c2];
    aux[4 * (int64_t)t + 3] = T(0.5) * P.gamma * F.curv[i1(d, i, j, k)] * F.cnorm[c2];
}

// ---------------------------------------------------------------------------------------
// Initial state (src/Init_multiphase.cpp:299-443)
// ---------------------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(256) void k_init_phi(Fields<T> F, int option, double z0, double phi_inlet, int open_z,
                                                  double sa_target, uint64_t seed) {
    const Dims& d = F.d;
    const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (t >= d.S4) return;
    const int i = (int)(t % d.X4) - 3, j = (int)((t / d.X4) % d.Y4) - 3, k = (int)(t / ((int64_t)d.X4 * d.Y4)) - 3;
    double v = 0.0;
    if (i >= 0 && j >= 0 && k >= 0 && i <= d.nx + 1 && j <= d.ny + 1 && k <= d.nz + 1) {
        const double x = i, y = j, z = k;
        if (option == 1) v = (z <= z0) ? 1.0 : -1.0;
        else if (option == 2) v = (z <= z0) ? -1.0 : 1.0;
        else if (option == 6) v = node_uniform(seed, (uint64_t)i1(d, i, j, k)) > sa_target ? -1.0 : 1.0;
        else {
            const double cx = (option == 5) ? (double)(d.nx + 1) * 0.5 : (double)(d.nx + 1) * 0.0;
            const double dx = x - cx, dz = z - (double)(d.nz + 1) * 0.5, dy = y - (double)(d.ny + 1) * 0.5;
            const bool inside = dx * dx + dz * dz + dy * dy <= z0 * z0;
            v = (option == 4) ? (inside ? -1.0 : 1.0) : (inside ? 1.0 : -1.0);
        }
    }
    if (open_z && (double)k <= 0) v = phi_inlet;
    F.phi[t] = (T)v;
}

template <typename T, bool RELAXED>
__global__ __launch_bounds__(256) void k_init_eq(Fields<T> F, const double* rho, const double* u, const double* v,
                                                 const double* w, double rho_ref) {
    const Dims& d = F.d;
    const int t = blockIdx.x * 256 + threadIdx.x;
    if (t >= d.S1) return;
    const int i = t % d.X1, j = (t / d.X1) % d.Y1, k = t / (d.X1 * d.Y1);
    if (ShiftedStorage<T>::value) {
        // colour 1 and the total's deviation from the rest state at rho_ref, in FP64
        const double uu = u ? u[t] : 0.0, vv = v ? v[t] : 0.0, ww = w ? w[t] : 0.0, rr = rho ? rho[t] : 1.0;
        const double usq = uu * uu + vv * vv + ww * ww;
        const double r1 = rr * (1.0 + (double)F.phi[i4(d, i, j, k)]) * 0.5;
        const double cu[19] = {0.0, uu, -uu, vv, -vv, ww, -ww, uu + vv, -uu + vv, uu - vv, -uu - vv,
                               uu + ww, -uu + ww, uu - ww, -uu - ww, vv + ww, -vv + ww, vv - ww, -vv - ww};
#pragma unroll
        for (int q = 0; q < 19; ++q) {
            const double wqq = q == 0 ? 1.0 / 3.0 : (q <= 6 ? 1.0 / 18.0 : 1.0 / 36.0);
            const double rqq = RELAXED ? rest_weight(q) : wqq;
            const double shape = wqq * (3.0 * cu[q] + 4.5 * cu[q] * cu[q] - 1.5 * usq);
            F.pdf[t + d.S1 * q] = (T)(r1 * rqq + r1 * shape);
            F.pdf[t + d.S1 * (q + 19)] = (T)((rr - rho_ref) * rqq + rho_ref * (rqq - rest_weight(q)) + rr * shape);
        }
        return;
    }
    const T uu = u ? (T)u[t] : T(0), vv = v ? (T)v[t] : T(0), ww = w ? (T)w[t] : T(0), rr = rho ? (T)rho[t] : T(1);
    const T usq = uu * uu + vv * vv + ww * ww;
    const T ph = F.phi[i4(d, i, j, k)];
    const T rc[2] = {rr * (T(1.0) + ph) * T(0.5), rr * (T(1.0) - ph) * T(0.5)};
    const T cu[19] = {T(0), uu, -uu, vv, -vv, ww, -ww, uu + vv, -uu + vv, uu - vv, -uu - vv,
                      uu + ww, -uu + ww, uu - ww, -uu - ww, vv + ww, -vv + ww, vv - ww, -vv - ww};
#pragma unroll
    for (int g = 0; g < 2; ++g) {
        const T r = rc[g];
        F.pdf[t + d.S1 * (19 * g)] = r * (RELAXED ? W<T>::r0 : W<T>::w0) + r * W<T>::w0 * (T(-1.5) * usq);
#pragma unroll
        for (int q = 1; q < 19; ++q) {
            const T wqq = q <= 6 ? W<T>::w1 : W<T>::w2;
            const T rqq = RELAXED ? (q <= 6 ? W<T>::r1 : W<T>::r2) : wqq;
            F.pdf[t + d.S1 * (q + 19 * g)] = r * rqq + r * wqq * (T(3.0) * cu[q] + T(4.5) * cu[q] * cu[q] - T(1.5) * usq);
        }
    }
}


// ---------------------------------------------------------------------------------------
// compute_macro_vars (src/Misc.cpp:222-274) and the monitor sums (src/Monitor.cpp:17-80),
// over the fluid list (solid nodes contribute exactly zero in the reference).
// ---------------------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ void macro_node(const Sparse<T>& S, const StepConst<T>& P, int t, double& rho, double& u,
                                           double& v, double& w) {
    constexpr bool SH = ShiftedStorage<T>::value;
    T ft[19];
#pragma unroll
    for (int q = 0; q < 19; ++q) ft[q] = SH ? S.f[S.at(q + 19, t)] : S.f[S.at(q, t)] + S.f[S.at(q + 19, t)];
    T r = ft[0];
#pragma unroll
    for (int q = 1; q < 19; ++q) r = r + ft[q];
    const T* a = S.aux + 4 * (int64_t)t;
    const T tmp = a[3];
    const T fx = tmp * a[0], fy = tmp * a[1], fz = tmp * a[2] + P.force_z;
    rho = SH ? (double)P.rho_ref + (double)r : (double)r;
    u = (double)(ft[1] - ft[2] + ft[7] - ft[8] + ft[9] - ft[10] + ft[11] - ft[12] + ft[13] - ft[14] - T(0.5) * fx);
    v = (double)(ft[3] - ft[4] + ft[7] + ft[8] - ft[9] - ft[10] + ft[15] - ft[16] + ft[17] - ft[18] - T(0.5) * fy);
    w = (double)(ft[5] - ft[6] + ft[11] + ft[12] - ft[13] - ft[14] + ft[15] + ft[16] - ft[17] - ft[18] - T(0.5) * fz);
}

template <typename T>
__global__ __launch_bounds__(256) void k_macro(Sparse<T> S, StepConst<T> P, Dims d, double* rho, double* u, double* v,
                                               double* w) {
    const int t = blockIdx.x * 256 + threadIdx.x;
    if (t >= S.nf) return;
    int i, j, k;
    unpack_ijk(S.list[t], i, j, k);
    const int c = i1(d, i, j, k);
    double r, uu, vv, ww;
    macro_node(S, P, t, r, uu, vv, ww);
    if (rho) rho[c] = r;
    if (u) u[c] = uu;
    if (v) v[c] = vv;
    if (w) w[c] = ww;
}

// Partial monitor sums: grid (nsplit, nz); block (b, k) reduces its share of the fluid
// nodes of slice k (list entries zstart[k-1] .. zstart[k]) into 10 FP64 values:
// fl1 fl2 vol1 vol2 mass1 mass2 pre umax2 usq1 usq2.
constexpr int MON_VALUES = 10;
template <typename T>
__global__ __launch_bounds__(256) void k_monitor_partial(Sparse<T> S, StepConst<T> P, const T* __restrict__ phi, Dims d,
                                                         const int32_t* __restrict__ zstart, double* __restrict__ part) {
    const int k = blockIdx.y;
    const int beg = zstart[k], end = zstart[k + 1];
    const int chunk = (end - beg + gridDim.x - 1) / gridDim.x;
    const int b0 = beg + blockIdx.x * chunk, b1 = min(end, b0 + chunk);
    double acc[MON_VALUES];
#pragma unroll
    for (int m = 0; m < MON_VALUES; ++m) acc[m] = 0.0;
    for (int t = b0 + threadIdx.x; t < b1; t += 256) {
        double r, u, v, w;
        macro_node(S, P, t, r, u, v, w);
        int i, j, kk;
        unpack_ijk(S.list[t], i, j, kk);
        const double ph = (double)phi[i4(d, i, j, kk)];
        // per-node terms in the reference's operand order (src/Monitor.cpp:60-67)
        acc[0] += w * 0.5 * (1.0 + ph);
        acc[1] += w * 0.5 * (1.0 - ph);
        acc[2] += 0.5 * (1.0 + ph);
        acc[3] += 0.5 * (1.0 - ph);
        acc[4] += r * 0.5 * (1.0 + ph);
        acc[5] += r * 0.5 * (1.0 - ph);
        acc[6] += r;
        const double u2 = u * u + v * v + w * w;
        acc[7] = fmax(acc[7], u2);
        if (ph > 0.999) acc[8] += u2;
        else if (ph < -0.999) acc[9] += u2;
    }
    double s;
    if (block_tree<MON_VALUES, 256, 7>(acc, s))
        part[((int64_t)blockIdx.y * gridDim.x + blockIdx.x) * MON_VALUES + threadIdx.x] = s;
}

// Combine the partials of each slice in block order: out[m * nz + k].
template <typename T>
__global__ void k_monitor_combine(const double* __restrict__ part, int nsplit, int nz, double* __restrict__ out) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nz) return;
    for (int m = 0; m < MON_VALUES; ++m) {
        double s = part[((int64_t)k * nsplit) * MON_VALUES + m];
        for (int b = 1; b < nsplit; ++b) {
            const double x = part[((int64_t)k * nsplit + b) * MON_VALUES + m];
            s = (m == 7) ? fmax(s, x) : s + x;
        }
        out[m * nz + k] = s;
    }
}

// Relative change of the velocity field since the last call (steady criterion of the warm
// start): per block, FP64 sums of |u - u_prev|^2 and |u|^2 over its fluid nodes, and u_prev
// replaced by u; k_sum_pairs then adds the block sums in block order (deterministic).
template <typename T>
__global__ __launch_bounds__(256) void k_velocity_change(Sparse<T> S, StepConst<T> P, double* __restrict__ prev,
                                                         double* __restrict__ part) {
    const int t = blockIdx.x * 256 + threadIdx.x;
    double d2 = 0.0, u2 = 0.0;
    if (t < S.nf) {
        double r, u, v, w;
        macro_node(S, P, t, r, u, v, w);
        const double du = u - prev[t], dv = v - prev[S.nf + t], dw = w - prev[2 * (int64_t)S.nf + t];
        d2 = du * du + dv * dv + dw * dw;
        u2 = u * u + v * v + w * w;
        prev[t] = u;
        prev[S.nf + t] = v;
        prev[2 * (int64_t)S.nf + t] = w;
    }
    double acc[2] = {d2, u2}, s;
    if (block_tree<2, 256>(acc, s)) part[2 * (int64_t)blockIdx.x + threadIdx.x] = s;
}

// out[0..1] = sums of the n pairs in part; one block, each thread a fixed stride.
template <typename T>
__global__ __launch_bounds__(256) void k_sum_pairs(const double* __restrict__ part, int n, double* __restrict__ out) {
    double a = 0.0, b = 0.0;
    for (int i = threadIdx.x; i < n; i += 256) {
        a += part[2 * (int64_t)i];
        b += part[2 * (int64_t)i + 1];
    }
    double acc[2] = {a, b}, s;
    if (block_tree<2, 256>(acc, s)) out[threadIdx.x] = s;
}

}  // namespace dev
}  // namespace mflbm
