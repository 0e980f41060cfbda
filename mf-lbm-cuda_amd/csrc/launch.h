// Host-side launchers of the step kernels, one explicit instantiation per precision
// (kernels_f64.hip, kernels_f32.hip).  All launches are asynchronous on `s`.
#pragma once
#include <hip/hip_runtime.h>
#include "lattice.h"

namespace mflbm {

template <typename T>
struct Launch {
    static void collide(const Sparse<T>& S, T* phi, const Dims& d, const StepConst<T>& P, bool odd, hipStream_t s);
    static void periodic_pdf(const Sparse<T>& S, const int32_t* pairs, int npairs, hipStream_t s);
    static void periodic_phi(T* phi, const Dims& d, int kper, int jper, hipStream_t s);
    static void open_bc(const Sparse<T>& S, const Fields<T>& F, const StepConst<T>& P, const int32_t* zmap, int in_bc,
                        int out_bc, const T* w_in, T* conv, bool after_odd, hipStream_t s);
    static void porous_plate(const Sparse<T>& S, const Dims& d, const StepConst<T>& P, const int32_t* pmap, int zp,
                             int blocked, bool after_odd, hipStream_t s);
    static void conv_init(const Fields<T>& F, T rho_ref, T* conv, hipStream_t s);
    static void phi_extrap(const Fields<T>& F, const NodeLists& L, hipStream_t s);
    static void normals(const Fields<T>& F, const StepConst<T>& P, const NodeLists& L, hipStream_t s);
    static void normal_extrap(const Fields<T>& F, const NodeLists& L, hipStream_t s);
    static void curvature(const Fields<T>& F, const StepConst<T>& P, const NodeLists& L, T* aux, hipStream_t s);
    static void gradient(const GradArgs<T>& A, const StepConst<T>& P, hipStream_t s);
    static void aux_refresh(const Fields<T>& F, const StepConst<T>& P, const NodeLists& L, T* aux, hipStream_t s);
    static void cells(const Sparse<T>& S, T* dense, const Dims& d, const int32_t* ext_node, const int32_t* ext_off,
                      int n_ext, bool to_dense, hipStream_t s);
    static void init_phi(const Fields<T>& F, int option, double z0, double phi_inlet, int open_z, double sa_target,
                         uint64_t seed, hipStream_t s);
    static void init_eq(const Fields<T>& F, const double* rho, const double* u, const double* v, const double* w,
                        bool relaxed, double rho_ref, hipStream_t s);
    static void macro(const Sparse<T>& S, const StepConst<T>& P, const Dims& d, double* rho, double* u, double* v,
                      double* w, hipStream_t s);
    // relative velocity change since the last call; part: 2 * ceil(nf / 256) doubles, out: 2
    static void velocity_change(const Sparse<T>& S, const StepConst<T>& P, double* prev, double* part, double* out,
                                hipStream_t s);
    static void monitor(const Sparse<T>& S, const StepConst<T>& P, const T* phi, const Dims& d, const int32_t* zstart,
                        int nsplit, double* part, double* out, hipStream_t s);
    static void to_real(const double* src, T* dst, int64_t n, hipStream_t s);
    // dst[4 i .. 4 i + 3] = (x[i], y[i], z[i], 0)
    static void pack4(const T* x, const T* y, const T* z, T* dst, int64_t n, hipStream_t s);
};

extern template struct Launch<double>;
extern template struct Launch<float>;

}  // namespace mflbm
