// Definitions of the Launch<T> members; included by kernels_f64.hip and kernels_f32.hip.
#pragma once
#include "kernels.inl"
#include "gradient.inl"
#include "launch.h"

namespace mflbm {

namespace {
inline unsigned blocks(int64_t n, int b = 256) { return (unsigned)((n + b - 1) / b); }

template <typename T>
__global__ __launch_bounds__(256) void k_pack4(const T* __restrict__ x, const T* __restrict__ y, const T* __restrict__ z,
                                               T* __restrict__ dst, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    dst[4 * i] = x[i];
    dst[4 * i + 1] = y[i];
    dst[4 * i + 2] = z[i];
    dst[4 * i + 3] = T(0);
}

template <typename T>
__global__ __launch_bounds__(256) void k_to_real(const double* __restrict__ src, T* __restrict__ dst, int64_t n) {
    const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (t < n) dst[t] = (T)src[t];
}
}  // namespace

template <typename T>
void Launch<T>::collide(const Sparse<T>& S, T* phi, const Dims& d, const StepConst<T>& P, bool odd, hipStream_t s) {
    if (S.nf <= 0) return;
    constexpr int b = MFLBM_COLLIDE_THREADS;
    if (P.mrt == 2) {
        if (odd) dev::k_collide<T, true><<<blocks(S.nf, b), b, 0, s>>>(S, phi, d, P);
        else dev::k_collide<T, false><<<blocks(S.nf, b), b, 0, s>>>(S, phi, d, P);
    } else {  // collision models 1, 3, 4
        if (odd) dev::k_collide<T, true, false><<<blocks(S.nf, b), b, 0, s>>>(S, phi, d, P);
        else dev::k_collide<T, false, false><<<blocks(S.nf, b), b, 0, s>>>(S, phi, d, P);
    }
}

template <typename T>
void Launch<T>::periodic_pdf(const Sparse<T>& S, const int32_t* pairs, int npairs, hipStream_t s) {
    if (npairs > 0) dev::k_periodic_pdf<T><<<blocks(npairs), 256, 0, s>>>(S, pairs, npairs);
}

template <typename T>
void Launch<T>::periodic_phi(T* phi, const Dims& d, int kper, int jper, hipStream_t s) {
    const int64_t n = (kper ? (int64_t)d.nx * d.ny : 0) + (jper ? (int64_t)d.nx * d.nz : 0) + ((kper && jper) ? d.nx : 0);
    if (n > 0) dev::k_periodic_phi<T><<<blocks(n), 256, 0, s>>>(phi, d, kper, jper);
}

template <typename T>
void Launch<T>::open_bc(const Sparse<T>& S, const Fields<T>& F, const StepConst<T>& P, const int32_t* zmap, int in_bc,
                        int out_bc, const T* w_in, T* conv, bool after_odd, hipStream_t s) {
    // inlet and outlet on separate threads when both are set and their planes are apart
    const int split = in_bc && out_bc && F.d.nz >= 6;
    const int n = (split ? 2 : 1) * F.d.nx * F.d.ny;
    if (after_odd) dev::k_open_bc<T, true><<<blocks(n), 256, 0, s>>>(S, F, P, zmap, in_bc, out_bc, w_in, conv, split);
    else dev::k_open_bc<T, false><<<blocks(n), 256, 0, s>>>(S, F, P, zmap, in_bc, out_bc, w_in, conv, split);
}

template <typename T>
void Launch<T>::porous_plate(const Sparse<T>& S, const Dims& d, const StepConst<T>& P, const int32_t* pmap, int zp,
                             int blocked, bool after_odd, hipStream_t s) {
    const int n = d.nx * d.ny;
    if (after_odd) dev::k_porous_plate<T, true><<<blocks(n), 256, 0, s>>>(S, d, P, pmap, zp, blocked);
    else dev::k_porous_plate<T, false><<<blocks(n), 256, 0, s>>>(S, d, P, pmap, zp, blocked);
}

template <typename T>
void Launch<T>::conv_init(const Fields<T>& F, T rho_ref, T* conv, hipStream_t s) {
    dev::k_conv_init<T><<<blocks((int64_t)F.d.X1 * F.d.Y1), 256, 0, s>>>(F, rho_ref, conv);
}

template <typename T>
void Launch<T>::phi_extrap(const Fields<T>& F, const NodeLists& L, hipStream_t s) {
    if (L.n_solid_bnd > 0) dev::k_phi_extrap<T><<<blocks(L.n_solid_bnd), 256, 0, s>>>(F, L.solid_bnd, L.n_solid_bnd);
}

template <typename T>
void Launch<T>::normals(const Fields<T>& F, const StepConst<T>& P, const NodeLists& L, hipStream_t s) {
    if (L.n_fluid_ext > 0) dev::k_normals<T><<<blocks(L.n_fluid_ext), 256, 0, s>>>(F, P, L.fluid_ext, L.n_fluid_ext);
}

template <typename T>
void Launch<T>::normal_extrap(const Fields<T>& F, const NodeLists& L, hipStream_t s) {
    if (L.n_solid_bnd > 0) dev::k_normal_extrap<T><<<blocks(L.n_solid_bnd), 256, 0, s>>>(F, L.solid_bnd, L.n_solid_bnd);
}

template <typename T>
void Launch<T>::curvature(const Fields<T>& F, const StepConst<T>& P, const NodeLists& L, T* aux, hipStream_t s) {
    if (L.n_fluid > 0) dev::k_curvature<T><<<blocks(L.n_fluid), 256, 0, s>>>(F, P, L.fluid, L.n_fluid, aux);
}

template <typename T>
void Launch<T>::gradient(const GradArgs<T>& A, const StepConst<T>& P, hipStream_t s) {
    if (A.shape == GRAD_SHAPE_TILES) {
        constexpr int NX = GradTiles<T>::N[0], NY = GradTiles<T>::N[1], NZ = GradTiles<T>::N[2];
        constexpr int FX = GradTiles<T>::F[0], FY = GradTiles<T>::F[1], FZ = GradTiles<T>::F[2];
        GradArgs<T> B = A;
        // normals over [-1, n+2]^3, forces over [1, n]^3
        B.tiles_x = (A.d.nx + 4 + NX - 1) / NX;
        B.tiles_y = (A.d.ny + 4 + NY - 1) / NY;
        B.tiles_z = (A.d.nz + 4 + NZ - 1) / NZ;
        const unsigned gn = (unsigned)(B.tiles_x * B.tiles_y * B.tiles_z);
        dev::k_grad_normals<T, NX, NY, NZ><<<gn, (NX + 4) * (NY + 4), 0, s>>>(B, P);
        B.tiles_x = (A.d.nx + FX - 1) / FX;
        B.tiles_y = (A.d.ny + FY - 1) / FY;
        B.tiles_z = (A.d.nz + FZ - 1) / FZ;
        const unsigned gf = (unsigned)(B.tiles_x * B.tiles_y * B.tiles_z);
        dev::k_grad_force<T, FX, FY, FZ><<<gf, (FX + 4) * (FY + 4), 0, s>>>(B, P);
        return;
    }
    const int chunks = (A.d.nz + A.kchunk - 1) / A.kchunk;
    const unsigned grid = (unsigned)(A.tiles_x * A.tiles_y * chunks);
    static_assert(GRAD_NSHAPES == 5, "one case per shape");
    switch (A.shape) {
#define MFLBM_GRAD_SHAPE(n)                                                                                  \
    case n:                                                                                                  \
        dev::k_gradient<T, GRAD_SHAPES[n].tx, GRAD_SHAPES[n].ty, GRAD_SHAPES[n].threads>                     \
            <<<grid, GRAD_SHAPES[n].threads, 0, s>>>(A, P);                                                  \
        break;
        MFLBM_GRAD_SHAPE(0)
        MFLBM_GRAD_SHAPE(1)
        MFLBM_GRAD_SHAPE(2)
        MFLBM_GRAD_SHAPE(3)
        MFLBM_GRAD_SHAPE(4)
#undef MFLBM_GRAD_SHAPE
        default: break;
    }
}

template <typename T>
void Launch<T>::velocity_change(const Sparse<T>& S, const StepConst<T>& P, double* prev, double* part, double* out,
                                hipStream_t s) {
    const int nb = (int)blocks(S.nf);
    if (nb > 0) dev::k_velocity_change<T><<<nb, 256, 0, s>>>(S, P, prev, part);
    dev::k_sum_pairs<T><<<1, 256, 0, s>>>(part, nb, out);
}

template <typename T>
void Launch<T>::aux_refresh(const Fields<T>& F, const StepConst<T>& P, const NodeLists& L, T* aux, hipStream_t s) {
    if (L.n_fluid > 0) dev::k_aux_refresh<T><<<blocks(L.n_fluid), 256, 0, s>>>(F, P, L.fluid, L.n_fluid, aux);
}

template <typename T>
void Launch<T>::cells(const Sparse<T>& S, T* dense, const Dims& d, const int32_t* ext_node, const int32_t* ext_off,
                      int n_ext, bool to_dense, hipStream_t s) {
    if (to_dense) {
        if (S.nf > 0) dev::k_fluid_cells<T, true><<<blocks(S.nf), 256, 0, s>>>(S, dense, d);
        if (n_ext > 0) dev::k_extra_cells<T, true><<<blocks(n_ext), 256, 0, s>>>(S, dense, d, ext_node, ext_off);
    } else {
        if (S.nf > 0) dev::k_fluid_cells<T, false><<<blocks(S.nf), 256, 0, s>>>(S, dense, d);
        if (n_ext > 0) dev::k_extra_cells<T, false><<<blocks(n_ext), 256, 0, s>>>(S, dense, d, ext_node, ext_off);
    }
}

template <typename T>
void Launch<T>::init_phi(const Fields<T>& F, int option, double z0, double phi_inlet, int open_z, double sa_target,
                         uint64_t seed, hipStream_t s) {
    dev::k_init_phi<T><<<blocks(F.d.S4), 256, 0, s>>>(F, option, z0, phi_inlet, open_z, sa_target, seed);
}

template <typename T>
void Launch<T>::init_eq(const Fields<T>& F, const double* rho, const double* u, const double* v, const double* w,
                        bool relaxed, double rho_ref, hipStream_t s) {
    if (relaxed) dev::k_init_eq<T, true><<<blocks(F.d.S1), 256, 0, s>>>(F, rho, u, v, w, rho_ref);
    else dev::k_init_eq<T, false><<<blocks(F.d.S1), 256, 0, s>>>(F, rho, u, v, w, rho_ref);
}

template <typename T>
void Launch<T>::macro(const Sparse<T>& S, const StepConst<T>& P, const Dims& d, double* rho, double* u, double* v,
                      double* w, hipStream_t s) {
    if (S.nf > 0) dev::k_macro<T><<<blocks(S.nf), 256, 0, s>>>(S, P, d, rho, u, v, w);
}

template <typename T>
void Launch<T>::monitor(const Sparse<T>& S, const StepConst<T>& P, const T* phi, const Dims& d, const int32_t* zstart,
                        int nsplit, double* part, double* out, hipStream_t s) {
    dim3 grid(nsplit, d.nz);
    dev::k_monitor_partial<T><<<grid, 256, 0, s>>>(S, P, phi, d, zstart, part);
    dev::k_monitor_combine<T><<<blocks(d.nz, 64), 64, 0, s>>>(part, nsplit, d.nz, out);
}

template <typename T>
void Launch<T>::pack4(const T* x, const T* y, const T* z, T* dst, int64_t n, hipStream_t s) {
    k_pack4<T><<<blocks(n), 256, 0, s>>>(x, y, z, dst, n);
}

template <typename T>
void Launch<T>::to_real(const double* src, T* dst, int64_t n, hipStream_t s) {
    k_to_real<T><<<blocks(n), 256, 0, s>>>(src, dst, n);
}

}  // namespace mflbm
