// Open-boundary column functions: Zou-He pressure inlet (k = 1) and outlet (k = nz) for
// one (i, j) column (src/main_iteration_GPU.cu:1085-1241 inlet, 1363-1524 outlet), the
// velocity inlet, the convective outlet and the porous plate.
//
// Written against three accessors so the same body serves the device kernel and the
// host-side pass that records which population slots the boundary touches:
//   f(a, b, c, e, g) -> T&   population slot e of colour g at node (a, b, c)
//   wall(a, b, c)    -> int  solid flag
//   ph(a, b, c)      -> T&   order parameter
// The sums follow the slot order of each reference variant; the orders differ between the
// before-odd and after-odd kernels and so does the rounding.  No branch depends on a
// population value, so a recording pass sees every access.
#pragma once
#include "lattice.h"

namespace mflbm {

template <typename T, bool AFTER_ODD, class Acc, class Wall, class Phi>
MFLBM_HD void zou_he_column(const StepConst<T>& P, int i, int j, int nz, int do_in, int do_out, Acc&& f, Wall&& wall,
                            Phi&& ph) {
    const T c3 = T(0.333333333333333333), c6 = T(0.166666666666666667);
    if (do_in) {
        const int w = wall(i, j, 1);
        T& p0 = ph(i, j, 0);
        p0 = P.phi_inlet * (1 - w) + p0 * w;
        const T pv = p0;
        ph(i, j, -1) = pv;
        ph(i, j, -2) = pv;
        ph(i, j, -3) = pv;
        T rT[2];
        rT[1] = P.rho_in;
        rT[0] = P.rho_in * P.sa_inject;
        rT[1] = rT[1] - rT[0];
        for (int g = 0; g < 2; ++g) {
            if (!AFTER_ODD) {
                T tt = (rT[g] - (f(i, j, 1, 0, g) + f(i - 1, j, 1, 1, g) + f(i + 1, j, 1, 2, g) + f(i, j - 1, 1, 3, g) +
                                 f(i, j + 1, 1, 4, g) + f(i - 1, j - 1, 1, 7, g) + f(i + 1, j - 1, 1, 8, g) +
                                 f(i - 1, j + 1, 1, 9, g) + f(i + 1, j + 1, 1, 10, g) +
                                 T(2.0) * (f(i, j, 2, 6, g) + f(i + 1, j, 2, 14, g) + f(i - 1, j, 2, 13, g) +
                                           f(i, j + 1, 2, 18, g) + f(i, j - 1, 2, 17, g)))) * P.relaxation;
                T tnx = T(0.5) * (f(i - 1, j, 1, 1, g) + f(i - 1, j - 1, 1, 7, g) + f(i - 1, j + 1, 1, 9, g) -
                                  (f(i + 1, j, 1, 2, g) + f(i + 1, j - 1, 1, 8, g) + f(i + 1, j + 1, 1, 10, g)));
                T tny = T(0.5) * (f(i, j - 1, 1, 3, g) + f(i - 1, j - 1, 1, 7, g) + f(i + 1, j - 1, 1, 8, g) -
                                  (f(i, j + 1, 1, 4, g) + f(i + 1, j + 1, 1, 10, g) + f(i - 1, j + 1, 1, 9, g)));
                f(i, j, 0, 5, g) = (f(i, j, 2, 6, g) + c3 * tt) * (1 - w) + f(i, j, 0, 5, g) * w;
                f(i - 1, j, 0, 11, g) = (f(i + 1, j, 2, 14, g) + c6 * tt - tnx) * (1 - w) + f(i - 1, j, 0, 11, g) * w;
                f(i + 1, j, 0, 12, g) = (f(i - 1, j, 2, 13, g) + c6 * tt + tnx) * (1 - w) + f(i + 1, j, 0, 12, g) * w;
                f(i, j - 1, 0, 15, g) = (f(i, j + 1, 2, 18, g) + c6 * tt - tny) * (1 - w) + f(i, j - 1, 0, 15, g) * w;
                f(i, j + 1, 0, 16, g) = (f(i, j - 1, 2, 17, g) + c6 * tt + tny) * (1 - w) + f(i, j + 1, 0, 16, g) * w;
            } else {
                T tt = (rT[g] - (f(i, j, 1, 0, g) + f(i, j, 1, 2, g) + f(i, j, 1, 1, g) + f(i, j, 1, 4, g) +
                                 f(i, j, 1, 3, g) + f(i, j, 1, 8, g) + f(i, j, 1, 7, g) + f(i, j, 1, 10, g) +
                                 f(i, j, 1, 9, g) +
                                 T(2.0) * (f(i, j, 1, 5, g) + f(i, j, 1, 11, g) + f(i, j, 1, 12, g) +
                                           f(i, j, 1, 15, g) + f(i, j, 1, 16, g)))) * P.relaxation;
                T tnx = T(0.5) * (f(i, j, 1, 2, g) + f(i, j, 1, 8, g) + f(i, j, 1, 10, g) -
                                  (f(i, j, 1, 1, g) + f(i, j, 1, 7, g) + f(i, j, 1, 9, g)));
                T tny = T(0.5) * (f(i, j, 1, 4, g) + f(i, j, 1, 9, g) + f(i, j, 1, 10, g) -
                                  (f(i, j, 1, 3, g) + f(i, j, 1, 8, g) + f(i, j, 1, 7, g)));
                f(i, j, 1, 6, g) = (f(i, j, 1, 5, g) + c3 * tt) * (1 - w) + f(i, j, 1, 6, g) * w;
                f(i, j, 1, 13, g) = (f(i, j, 1, 12, g) + c6 * tt + tnx) * (1 - w) + f(i, j, 1, 13, g) * w;
                f(i, j, 1, 14, g) = (f(i, j, 1, 11, g) + c6 * tt - tnx) * (1 - w) + f(i, j, 1, 14, g) * w;
                f(i, j, 1, 17, g) = (f(i, j, 1, 16, g) + c6 * tt + tny) * (1 - w) + f(i, j, 1, 17, g) * w;
                f(i, j, 1, 18, g) = (f(i, j, 1, 15, g) + c6 * tt - tny) * (1 - w) + f(i, j, 1, 18, g) * w;
            }
        }
    }
    if (do_out) {
        const int w = wall(i, j, nz);
        const T phv = ph(i, j, nz);
        for (int m = 1; m <= 4; ++m) ph(i, j, nz + m) = phv;
        T t1;
        if (!AFTER_ODD) {
            t1 = f(i, j, nz, 0, 0) + f(i - 1, j, nz, 1, 0) + f(i + 1, j, nz, 2, 0) + f(i, j - 1, nz, 3, 0) +
                 f(i, j + 1, nz, 4, 0) + f(i - 1, j - 1, nz, 7, 0) + f(i + 1, j - 1, nz, 8, 0) +
                 f(i - 1, j + 1, nz, 9, 0) + f(i + 1, j + 1, nz, 10, 0) +
                 T(2.0) * (f(i, j, nz - 1, 5, 0) + f(i - 1, j, nz - 1, 11, 0) + f(i + 1, j, nz - 1, 12, 0) +
                           f(i, j - 1, nz - 1, 15, 0) + f(i, j + 1, nz - 1, 16, 0)) +
                 f(i, j, nz, 0, 1) + f(i - 1, j, nz, 1, 1) + f(i + 1, j, nz, 2, 1) + f(i, j - 1, nz, 3, 1) +
                 f(i, j + 1, nz, 4, 1) + f(i - 1, j - 1, nz, 7, 1) + f(i + 1, j - 1, nz, 8, 1) +
                 f(i - 1, j + 1, nz, 9, 1) + f(i + 1, j + 1, nz, 10, 1) +
                 T(2.0) * (f(i, j, nz - 1, 5, 1) + f(i - 1, j, nz - 1, 11, 1) + f(i + 1, j, nz - 1, 12, 1) +
                           f(i, j - 1, nz - 1, 15, 1) + f(i, j + 1, nz - 1, 16, 1)) -
                 P.rho_out;
        } else {
            t1 = f(i, j, nz, 0, 0) + f(i, j, nz, 2, 0) + f(i, j, nz, 1, 0) + f(i, j, nz, 4, 0) + f(i, j, nz, 3, 0) +
                 f(i, j, nz, 8, 0) + f(i, j, nz, 7, 0) + f(i, j, nz, 10, 0) + f(i, j, nz, 9, 0) +
                 T(2.0) * (f(i, j, nz, 6, 0) + f(i, j, nz, 14, 0) + f(i, j, nz, 13, 0) + f(i, j, nz, 18, 0) +
                           f(i, j, nz, 17, 0)) +
                 f(i, j, nz, 0, 1) + f(i, j, nz, 2, 1) + f(i, j, nz, 1, 1) + f(i, j, nz, 4, 1) + f(i, j, nz, 3, 1) +
                 f(i, j, nz, 8, 1) + f(i, j, nz, 7, 1) + f(i, j, nz, 10, 1) + f(i, j, nz, 9, 1) +
                 T(2.0) * (f(i, j, nz, 6, 1) + f(i, j, nz, 14, 1) + f(i, j, nz, 13, 1) + f(i, j, nz, 18, 1) +
                           f(i, j, nz, 17, 1)) -
                 P.rho_out;
        }
        T t2 = t1 * T(0.5) * (T(1.0) - phv);
        t1 = t1 - t2;
        const T tt2[2] = {t1, t2};
        for (int g = 0; g < 2; ++g) {
            const T tt = tt2[g];
            if (!AFTER_ODD) {
                T tnx = T(0.5) * (f(i - 1, j, nz, 1, g) + f(i - 1, j - 1, nz, 7, g) + f(i - 1, j + 1, nz, 9, g) -
                                  (f(i + 1, j, nz, 2, g) + f(i + 1, j - 1, nz, 8, g) + f(i + 1, j + 1, nz, 10, g)));
                T tny = T(0.5) * (f(i, j - 1, nz, 3, g) + f(i - 1, j - 1, nz, 7, g) + f(i + 1, j - 1, nz, 8, g) -
                                  (f(i, j + 1, nz, 4, g) + f(i + 1, j + 1, nz, 10, g) + f(i - 1, j + 1, nz, 9, g)));
                f(i, j, nz + 1, 6, g) = (f(i, j, nz - 1, 5, g) - c3 * tt) * (1 - w) + f(i, j, nz + 1, 6, g) * w;
                f(i - 1, j, nz + 1, 13, g) = (f(i + 1, j, nz - 1, 12, g) - c6 * tt - tnx) * (1 - w) + f(i - 1, j, nz + 1, 13, g) * w;
                f(i + 1, j, nz + 1, 14, g) = (f(i - 1, j, nz - 1, 11, g) - c6 * tt + tnx) * (1 - w) + f(i + 1, j, nz + 1, 14, g) * w;
                f(i, j - 1, nz + 1, 17, g) = (f(i, j + 1, nz - 1, 16, g) - c6 * tt - tny) * (1 - w) + f(i, j - 1, nz + 1, 17, g) * w;
                f(i, j + 1, nz + 1, 18, g) = (f(i, j - 1, nz - 1, 15, g) - c6 * tt + tny) * (1 - w) + f(i, j + 1, nz + 1, 18, g) * w;
            } else {
                T tnx = T(0.5) * (f(i, j, nz, 2, g) + f(i, j, nz, 8, g) + f(i, j, nz, 10, g) -
                                  (f(i, j, nz, 1, g) + f(i, j, nz, 7, g) + f(i, j, nz, 9, g)));
                T tny = T(0.5) * (f(i, j, nz, 4, g) + f(i, j, nz, 10, g) + f(i, j, nz, 9, g) -
                                  (f(i, j, nz, 3, g) + f(i, j, nz, 7, g) + f(i, j, nz, 8, g)));
                f(i, j, nz, 5, g) = (f(i, j, nz, 6, g) - c3 * tt) * (1 - w) + f(i, j, nz, 5, g) * w;
                f(i, j, nz, 11, g) = (f(i, j, nz, 14, g) - c6 * tt + tnx) * (1 - w) + f(i, j, nz, 11, g) * w;
                f(i, j, nz, 12, g) = (f(i, j, nz, 13, g) - c6 * tt - tnx) * (1 - w) + f(i, j, nz, 12, g) * w;
                f(i, j, nz, 15, g) = (f(i, j, nz, 18, g) - c6 * tt + tny) * (1 - w) + f(i, j, nz, 15, g) * w;
                f(i, j, nz, 16, g) = (f(i, j, nz, 17, g) - c6 * tt - tny) * (1 - w) + f(i, j, nz, 16, g) * w;
            }
        }
    }
}

// Velocity inlet at k = 1 by bounce-back with the injected momentum of the column's inlet
// velocity w_in, split between the colours by the injected saturation
// (src/main_iteration_GPU.cu:1008-1077, inlet_bounce_back_velocity_BC_{before,after}_odd).
template <typename T, bool AFTER_ODD, class Acc, class Wall, class Phi>
MFLBM_HD void velocity_inlet_column(const StepConst<T>& P, int i, int j, T w_in, Acc&& f, Wall&& wall, Phi&& ph) {
    const int w = wall(i, j, 1);
    T& p0 = ph(i, j, 0);
    p0 = P.phi_inlet * (1 - w) + p0 * w;
    const T pv = p0;
    ph(i, j, -1) = pv;
    ph(i, j, -2) = pv;
    ph(i, j, -3) = pv;
    T t2 = w_in * P.relaxation;
    const T t1 = t2 * P.sa_inject;  // fluid 1
    t2 = t2 - t1;                   // fluid 2
    const T a1 = T(6.0) * (T(1.0) / T(18.0)), a2 = T(6.0) * (T(1.0) / T(36.0));
    for (int g = 0; g < 2; ++g) {
        const T tg = g == 0 ? t1 : t2;
        if (!AFTER_ODD) {
            f(i, j, 0, 5, g) = (f(i, j, 1, 6, g) + a1 * tg) * (1 - w) + f(i, j, 0, 5, g) * w;
            f(i - 1, j, 0, 11, g) = (f(i, j, 1, 14, g) + a2 * tg) * (1 - w) + f(i - 1, j, 0, 11, g) * w;
            f(i + 1, j, 0, 12, g) = (f(i, j, 1, 13, g) + a2 * tg) * (1 - w) + f(i + 1, j, 0, 12, g) * w;
            f(i, j - 1, 0, 15, g) = (f(i, j, 1, 18, g) + a2 * tg) * (1 - w) + f(i, j - 1, 0, 15, g) * w;
            f(i, j + 1, 0, 16, g) = (f(i, j, 1, 17, g) + a2 * tg) * (1 - w) + f(i, j + 1, 0, 16, g) * w;
        } else {
            f(i, j, 1, 6, g) = (f(i, j, 0, 5, g) + a1 * tg) * (1 - w) + f(i, j, 1, 6, g) * w;
            f(i, j, 1, 13, g) = (f(i + 1, j, 0, 12, g) + a2 * tg) * (1 - w) + f(i, j, 1, 13, g) * w;
            f(i, j, 1, 14, g) = (f(i - 1, j, 0, 11, g) + a2 * tg) * (1 - w) + f(i, j, 1, 14, g) * w;
            f(i, j, 1, 17, g) = (f(i, j + 1, 0, 16, g) + a2 * tg) * (1 - w) + f(i, j, 1, 17, g) * w;
            f(i, j, 1, 18, g) = (f(i, j - 1, 0, 15, g) + a2 * tg) * (1 - w) + f(i, j, 1, 18, g) * w;
        }
    }
}

// Convective outlet at k = nz: the unknown populations and the ghost phi follow
// d/dt + u d/dz = 0 with u = P.uin_avg, from the values stored for the column at the last
// step (cs(e, g) populations, cphi() phi), which are then replaced
// (src/main_iteration_GPU.cu:1245-1357, outlet_convective_BC_{before,after}_odd).
template <typename T, bool AFTER_ODD, class Acc, class Wall, class Phi, class CS, class CP>
MFLBM_HD void convective_outlet_column(const StepConst<T>& P, int i, int j, int nz, Acc&& f, Wall&& wall, Phi&& ph,
                                       CS&& cs, CP&& cphi) {
    const T u = P.uin_avg;
    const T tmp = T(1.0) / (T(1.0) + u);
    const int w = wall(i, j, nz);
    T& pg = ph(i, j, nz + 1);
    pg = ((cphi() + u * ph(i, j, nz)) * tmp) * (1 - w) + pg * w;
    cphi() = pg;
    const T pv = pg;
    ph(i, j, nz + 2) = pv;
    ph(i, j, nz + 3) = pv;
    ph(i, j, nz + 4) = pv;
    for (int g = 0; g < 2; ++g) {
        if (!AFTER_ODD) {
            f(i, j, nz + 1, 6, g) = ((cs(6, g) + u * f(i, j, nz, 6, g)) * tmp) * (1 - w) + f(i, j, nz + 1, 6, g) * w;
            f(i - 1, j, nz + 1, 13, g) =
                ((cs(13, g) + u * f(i - 1, j, nz, 13, g)) * tmp) * (1 - w) + f(i - 1, j, nz + 1, 13, g) * w;
            f(i + 1, j, nz + 1, 14, g) =
                ((cs(14, g) + u * f(i + 1, j, nz, 14, g)) * tmp) * (1 - w) + f(i + 1, j, nz + 1, 14, g) * w;
            f(i, j - 1, nz + 1, 17, g) =
                ((cs(17, g) + u * f(i, j - 1, nz, 17, g)) * tmp) * (1 - w) + f(i, j - 1, nz + 1, 17, g) * w;
            f(i, j + 1, nz + 1, 18, g) =
                ((cs(18, g) + u * f(i, j + 1, nz, 18, g)) * tmp) * (1 - w) + f(i, j + 1, nz + 1, 18, g) * w;
        } else {
            f(i, j, nz, 5, g) = ((cs(6, g) + u * f(i, j, nz - 1, 5, g)) * tmp) * (1 - w) + f(i, j, nz, 5, g) * w;
            f(i, j, nz, 11, g) = ((cs(14, g) + u * f(i, j, nz - 1, 11, g)) * tmp) * (1 - w) + f(i, j, nz, 11, g) * w;
            f(i, j, nz, 12, g) = ((cs(13, g) + u * f(i, j, nz - 1, 12, g)) * tmp) * (1 - w) + f(i, j, nz, 12, g) * w;
            f(i, j, nz, 15, g) = ((cs(18, g) + u * f(i, j, nz - 1, 15, g)) * tmp) * (1 - w) + f(i, j, nz, 15, g) * w;
            f(i, j, nz, 16, g) = ((cs(17, g) + u * f(i, j, nz - 1, 16, g)) * tmp) * (1 - w) + f(i, j, nz, 16, g) * w;
        }
    }
    for (int g = 0; g < 2; ++g) {
        if (!AFTER_ODD) {
            cs(6, g) = f(i, j, nz + 1, 6, g);
            cs(13, g) = f(i - 1, j, nz + 1, 13, g);
            cs(14, g) = f(i + 1, j, nz + 1, 14, g);
            cs(17, g) = f(i, j - 1, nz + 1, 17, g);
            cs(18, g) = f(i, j + 1, nz + 1, 18, g);
        } else {
            cs(6, g) = f(i, j, nz, 5, g);
            cs(14, g) = f(i, j, nz, 11, g);
            cs(13, g) = f(i, j, nz, 12, g);
            cs(18, g) = f(i, j, nz, 15, g);
            cs(17, g) = f(i, j, nz, 16, g);
        }
    }
}

// Porous plate at k = zp that lets one colour through and reflects the other, `blocked`
// (0 = fluid 1, 1 = fluid 2) (src/main_iteration_GPU.cu:1744-1882, porous_plate_BC_*).
template <bool AFTER_ODD, class Acc>
MFLBM_HD void porous_plate_column(int i, int j, int zp, int blocked, Acc&& f) {
    const int b = blocked, c = 1 - blocked;
    if (!AFTER_ODD) {
        f(i, j, zp, 6, b) = f(i, j, zp - 1, 5, b);
        f(i - 1, j, zp, 13, b) = f(i, j, zp - 1, 12, b);
        f(i + 1, j, zp, 14, b) = f(i, j, zp - 1, 11, b);
        f(i, j - 1, zp, 17, b) = f(i, j, zp - 1, 16, b);
        f(i, j + 1, zp, 18, b) = f(i, j, zp - 1, 15, b);
        f(i, j, zp, 5, b) = f(i, j, zp + 1, 6, b);
        f(i + 1, j, zp, 12, b) = f(i, j, zp + 1, 13, b);
        f(i - 1, j, zp, 11, b) = f(i, j, zp + 1, 14, b);
        f(i, j + 1, zp, 16, b) = f(i, j, zp + 1, 17, b);
        f(i, j - 1, zp, 15, b) = f(i, j, zp + 1, 18, b);
        const int up[5] = {6, 13, 14, 17, 18}, dn[5] = {5, 12, 11, 16, 15};
        for (int m = 0; m < 5; ++m) f(i, j, zp, up[m], c) = f(i, j, zp + 1, up[m], c);
        for (int m = 0; m < 5; ++m) f(i, j, zp, dn[m], c) = f(i, j, zp - 1, dn[m], c);
    } else {
        f(i, j, zp - 1, 5, b) = f(i, j, zp, 6, b);
        f(i, j, zp - 1, 12, b) = f(i - 1, j, zp, 13, b);
        f(i, j, zp - 1, 11, b) = f(i + 1, j, zp, 14, b);
        f(i, j, zp - 1, 16, b) = f(i, j - 1, zp, 17, b);
        f(i, j, zp - 1, 15, b) = f(i, j + 1, zp, 18, b);
        f(i, j, zp + 1, 6, b) = f(i, j, zp, 5, b);
        f(i, j, zp + 1, 13, b) = f(i + 1, j, zp, 12, b);
        f(i, j, zp + 1, 14, b) = f(i - 1, j, zp, 11, b);
        f(i, j, zp + 1, 17, b) = f(i, j + 1, zp, 16, b);
        f(i, j, zp + 1, 18, b) = f(i, j - 1, zp, 15, b);
        const int up[5] = {6, 13, 14, 17, 18}, dn[5] = {5, 12, 11, 16, 15};
        for (int m = 0; m < 5; ++m) f(i, j, zp - 1, dn[m], c) = f(i, j, zp, dn[m], c);
        for (int m = 0; m < 5; ++m) f(i, j, zp + 1, up[m], c) = f(i, j, zp, up[m], c);
    }
}

// Population slots the periodic ghost copies move, in the reference's kernel order
// (Z faces, Y faces, ZY edges; src/main_iteration_GPU.cu:1529-1689).  cb(dst_node,
// src_node, slot) receives s1 node indices; a later call for the same destination
// overrides an earlier one, as the later reference kernel overwrites it.
template <class CB>
inline void periodic_pdf_copies(const Dims& d, int kper, int jper, bool odd, CB&& cb) {
    const int nx = d.nx, ny = d.ny, nz = d.nz;
    if (kper) {
        const int up[5] = {6, 14, 13, 18, 17}, dn[5] = {5, 11, 12, 15, 16};
        for (int j = 1; j <= ny; ++j)
            for (int i = 1; i <= nx; ++i)
                for (int m = 0; m < 5; ++m) {
                    if (!odd) {
                        cb(i1(d, i, j, nz + 1), i1(d, i, j, 1), up[m]);
                        cb(i1(d, i, j, 0), i1(d, i, j, nz), dn[m]);
                    } else {
                        cb(i1(d, i, j, 1), i1(d, i, j, nz + 1), up[m]);
                        cb(i1(d, i, j, nz), i1(d, i, j, 0), dn[m]);
                    }
                }
    }
    if (jper) {
        const int up[5] = {4, 10, 9, 16, 18}, dn[5] = {3, 7, 8, 15, 17};
        for (int k = 1; k <= nz; ++k)
            for (int i = 1; i <= nx; ++i)
                for (int m = 0; m < 5; ++m) {
                    if (!odd) {
                        cb(i1(d, i, ny + 1, k), i1(d, i, 1, k), up[m]);
                        cb(i1(d, i, 0, k), i1(d, i, ny, k), dn[m]);
                    } else {
                        cb(i1(d, i, 1, k), i1(d, i, ny + 1, k), up[m]);
                        cb(i1(d, i, ny, k), i1(d, i, 0, k), dn[m]);
                    }
                }
    }
    if (kper && jper) {
        for (int i = 1; i <= nx; ++i) {
            if (!odd) {
                cb(i1(d, i, ny + 1, nz + 1), i1(d, i, 1, 1), 18);
                cb(i1(d, i, ny + 1, 0), i1(d, i, 1, nz), 16);
                cb(i1(d, i, 0, nz + 1), i1(d, i, ny, 1), 17);
                cb(i1(d, i, 0, 0), i1(d, i, ny, nz), 15);
            } else {
                cb(i1(d, i, 1, 1), i1(d, i, ny + 1, nz + 1), 18);
                cb(i1(d, i, 1, nz), i1(d, i, ny + 1, 0), 16);
                cb(i1(d, i, ny, 1), i1(d, i, 0, nz + 1), 17);
                cb(i1(d, i, ny, nz), i1(d, i, 0, 0), 15);
            }
        }
    }
}

}  // namespace mflbm
