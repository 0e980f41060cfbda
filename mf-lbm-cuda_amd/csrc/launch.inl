// Definitions of the Launch<T> members; included by kernels_f64.hip and kernels_f32.hip.
#pragma once
#include "kernels.inl"
#include "launch.h"

namespace mflbm {

namespace {
inline unsigned blocks(int64_t n, int b = 256) { return (unsigned)((n + b - 1) / b); }

template <typename T>
__global__ __launch_bounds__(256) void k_to_real(const double* __restrict__ src, T* __restrict__ dst, int64_t n) {
    const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (t < n) dst[t] = (T)src[t];
}
}  // namespace

template <typename T>
void Launch<T>::collide(const Fields<T>& F, const StepConst<T>& P, const NodeLists& L, bool odd, hipStream_t s) {
    if (L.n_fluid <= 0) return;
    if (odd) dev::k_collide<T, true><<<blocks(L.n_fluid), 256, 0, s>>>(F, P, L.fluid, L.n_fluid);
    else dev::k_collide<T, false><<<blocks(L.n_fluid), 256, 0, s>>>(F, P, L.fluid, L.n_fluid);
}

template <typename T>
void Launch<T>::periodic(const Fields<T>& F, int kper, int jper, bool odd, hipStream_t s) {
    const int64_t n = (kper ? (int64_t)F.d.nx * F.d.ny : 0) + (jper ? (int64_t)F.d.nx * F.d.nz : 0) +
                      ((kper && jper) ? F.d.nx : 0);
    if (n == 0) return;
    if (odd) dev::k_periodic<T, true><<<blocks(n), 256, 0, s>>>(F, kper, jper);
    else dev::k_periodic<T, false><<<blocks(n), 256, 0, s>>>(F, kper, jper);
}

template <typename T>
void Launch<T>::zou_he(const Fields<T>& F, const StepConst<T>& P, int do_in, int do_out, bool after_odd, hipStream_t s) {
    const int64_t n = (int64_t)F.d.nx * F.d.ny;
    if (after_odd) dev::k_zou_he<T, true><<<blocks(n), 256, 0, s>>>(F, P, do_in, do_out);
    else dev::k_zou_he<T, false><<<blocks(n), 256, 0, s>>>(F, P, do_in, do_out);
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
void Launch<T>::curvature(const Fields<T>& F, const NodeLists& L, hipStream_t s) {
    if (L.n_fluid > 0) dev::k_curvature<T><<<blocks(L.n_fluid), 256, 0, s>>>(F, L.fluid, L.n_fluid);
}

template <typename T>
void Launch<T>::init_phi(const Fields<T>& F, int option, double z0, double phi_inlet, int open_z, hipStream_t s) {
    dev::k_init_phi<T><<<blocks(F.d.S4), 256, 0, s>>>(F, option, z0, phi_inlet, open_z);
}

template <typename T>
void Launch<T>::init_eq(const Fields<T>& F, const double* rho, const double* u, const double* v, const double* w,
                        hipStream_t s) {
    dev::k_init_eq<T><<<blocks(F.d.S1), 256, 0, s>>>(F, rho, u, v, w);
}

template <typename T>
void Launch<T>::macro(const Fields<T>& F, const StepConst<T>& P, double* rho, double* u, double* v, double* w,
                      hipStream_t s) {
    dev::k_macro<T><<<blocks((int64_t)F.d.nx * F.d.ny * F.d.nz), 256, 0, s>>>(F, P, rho, u, v, w);
}

template <typename T>
void Launch<T>::monitor(const Fields<T>& F, const StepConst<T>& P, int rows, int nsplit, double* part, double* out,
                        hipStream_t s) {
    dim3 grid(nsplit, F.d.nz);
    dev::k_monitor_partial<T><<<grid, 256, 0, s>>>(F, P, rows, part);
    dev::k_monitor_combine<T><<<blocks(F.d.nz, 64), 64, 0, s>>>(part, nsplit, F.d.nz, out);
}

template <typename T>
void Launch<T>::to_real(const double* src, T* dst, int64_t n, hipStream_t s) {
    k_to_real<T><<<blocks(n), 256, 0, s>>>(src, dst, n);
}

}  // namespace mflbm
