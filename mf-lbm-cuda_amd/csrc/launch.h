// Host-side launchers of the step kernels, one explicit instantiation per precision
// (kernels_f64.hip, kernels_f32.hip).  All launches are asynchronous on `s`.
#pragma once
#include <hip/hip_runtime.h>
#include "lattice.h"

namespace mflbm {

template <typename T>
struct Launch {
    static void collide(const Fields<T>& F, const StepConst<T>& P, const NodeLists& L, bool odd, hipStream_t s);
    static void periodic(const Fields<T>& F, int kper, int jper, bool odd, hipStream_t s);
    static void zou_he(const Fields<T>& F, const StepConst<T>& P, int do_in, int do_out, bool after_odd, hipStream_t s);
    static void phi_extrap(const Fields<T>& F, const NodeLists& L, hipStream_t s);
    static void normals(const Fields<T>& F, const StepConst<T>& P, const NodeLists& L, hipStream_t s);
    static void normal_extrap(const Fields<T>& F, const NodeLists& L, hipStream_t s);
    static void curvature(const Fields<T>& F, const NodeLists& L, hipStream_t s);
    static void init_phi(const Fields<T>& F, int option, double z0, double phi_inlet, int open_z, hipStream_t s);
    static void init_eq(const Fields<T>& F, const double* rho, const double* u, const double* v, const double* w,
                        hipStream_t s);
    static void macro(const Fields<T>& F, const StepConst<T>& P, double* rho, double* u, double* v, double* w,
                      hipStream_t s);
    static void monitor(const Fields<T>& F, const StepConst<T>& P, int rows, int nsplit, double* part, double* out,
                        hipStream_t s);
    static void to_real(const double* src, T* dst, int64_t n, hipStream_t s);
};

extern template struct Launch<double>;
extern template struct Launch<float>;

}  // namespace mflbm
