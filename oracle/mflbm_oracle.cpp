// =====================================================================================
//  MF-LBM colour-gradient step -- CPU FP64 ORACLE (test infrastructure only).
//
//  This file is the checker, never the product.  Only tests/, the smoke entry point and
//  the CPU-baseline leg of bench.py load it.  It restates, in plain OpenMP C++ and in
//  double precision, the arithmetic of the reference solver so that the HIP kernels can
//  be compared against it bit for bit.  Every routine cites the reference lines it
//  follows (paths relative to the reference repository root).
//
//  Floating-point contract: build with -O2 -ffp-contract=off and no fast-math, so that
//  every expression is evaluated as written (left-to-right association, one rounding
//  per operation, IEEE division and square root).  The expression order of the
//  reference is kept wherever it changes rounding; that is the only reason the long
//  sums below are spelled out term by term.
//
//  Pinning (how the restatement itself is checked, see tests/test_oracle_*.py):
//    collide_node   : independent matrix MRT + Guo forcing + LKR recolouring (numpy), for
//                     each of the four collision models; SRT at rest == BGK towards the
//                     collision's rest weights (4/21, 13/126, 1/63)
//    AA streaming   : two steps with zero-force rest populations vs explicit pull stream
//    gradient chain : ISO4 exact on linear/quadratic fields, sphere curvature ~ 2/R,
//                     wetting angle n.s_n = cos(theta') after the secant iteration
//    geometry       : brute-force walls_type classification; the smoothed indicator equals
//                     four passes of the separable (1,4,1)/6 filter along x, y and z; the
//                     solid normals equal an ISO8 gradient built by enumerating the
//                     integer offsets (reference form, linear gain 37/42); flat-wall
//                     normal length |nw|/(|nw|+FLT_EPSILON) in exact rationals; 45-degree
//                     slab and sphere normal directions
//    Zou-He BCs     : imposed per-colour density and zero transverse momentum
//    periodic copies: periodic run == run on a wrapped domain
//    monitor        : brute-force sums in numpy
// =====================================================================================
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <cfloat>
#include <vector>
#include <string>
#include <omp.h>

extern "C" {

// Parameters of one run.  Field meanings follow the reference globals
// (includes/Module.h, includes/Fluid_multiphase.h); derived values are passed in
// already derived (src/IO_multiphase.cpp:200-206, src/Init_multiphase.cpp:173-214).
typedef struct oracle_params {
    int64_t nx, ny, nz;
    int32_t iper, jper, kper;
    int32_t wall_x_min, wall_x_max, wall_y_min, wall_y_max, wall_z_min, wall_z_max;
    int32_t inlet_bc, outlet_bc;
    double lbm_gamma;   // surface tension
    double force_z;     // body force along z (0 with pressure BCs)
    double la_nui1;     // 1/nu1
    double la_nui2;     // 1/nu2
    double lbm_beta;    // recolouring beta
    double phi_inlet;   // 2*S_inj - 1
    double sa_inject;   // injected fluid-1 saturation
    double cos_theta;   // cos(180deg - theta)
    double rho_in;
    double rho_out;
    double relaxation;  // 1 in the reference (src/main.cpp:57)
    int32_t porous_plate;    // porous_plate_cmd: 0 none, 1 blocks fluid 1, 2 blocks fluid 2
    int32_t z_porous_plate;  // Z_porous_plate
    double uin_avg;          // capillary_number * surface_tension / nu1 (src/Init_multiphase.cpp:262)
    int32_t mrt;             // collision model (includes/preprocessor.h:3-4), 0 = 2
} oracle_params;

typedef struct oracle_monitor {
    // per-slice sums, length nz each (src/Monitor.cpp:52-80)
    double *fl1, *fl2, *vol1, *vol2, *mass1, *mass2, *pre;
    double umax2;       // max |u|^2 over fluid nodes (src/Monitor.cpp:34-50)
    double usq1, usq2;  // sum |u|^2 where phi > 0.999 / < -0.999
} oracle_monitor;

}  // extern "C"

namespace {

// ---- lattice tables (includes/Module.h:98-122) --------------------------------------
const int EX[19] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0};
const int EY[19] = {0, 0, 0, 1, -1, 0, 0, 1, 1, -1, -1, 0, 0, 0, 0, 1, -1, 1, -1};
const int EZ[19] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, 1, -1, -1, 1, 1, -1, -1};
const int OPP[19] = {0, 2, 1, 4, 3, 6, 5, 10, 9, 8, 7, 14, 13, 12, 11, 18, 17, 16, 15};

struct Consts {
    double w[19];
    double w1, w2, rk2;
    double r0, r1, r2;  // rest weights of the collision's equilibrium
    double pi;
    double iso4[2];
    double iso8[7];
    double c19, c2394, c252, c72;   // MRT back-transform coefficients
    double e2c1, e2c2, omega_xx;
    double eps;                     // FLT_EPSILON, also in double builds (includes/Module.h:10-16)
    Consts() {
        w[0] = 1.0 / 3.0;
        for (int q = 1; q <= 6; ++q) w[q] = 1.0 / 18.0;
        for (int q = 7; q <= 18; ++q) w[q] = 1.0 / 36.0;
        w1 = 1.0 / 18.0;
        w2 = 1.0 / 36.0;
        // rest state of the collision (e = -11 rho, e2 = e2c1 rho = 0): no reference counterpart
        r0 = 4.0 / 21.0; r1 = 13.0 / 126.0; r2 = 1.0 / 63.0;
        pi = 3.14159265358979323846;  // includes/Module.h:9
        rk2 = 1.0 / std::sqrt(2.0) / 36.0;           // includes/Fluid_multiphase.h:32
        iso4[0] = 1.0 / 6.0;  iso4[1] = 1.0 / 12.0;  // includes/Fluid_multiphase.h:34
        iso8[0] = 4.0 / 45.0; iso8[1] = 1.0 / 21.0; iso8[2] = 2.0 / 105.0; iso8[3] = 5.0 / 504.0;
        iso8[4] = 1.0 / 315.0; iso8[5] = 1.0 / 630.0; iso8[6] = 1.0 / 5040.0;  // :35
        c19 = 1.0 / 19.0; c2394 = 1.0 / 2394.0; c252 = 1.0 / 252.0; c72 = 1.0 / 72.0;
        e2c1 = 0.0; e2c2 = -475.0 / 63.0; omega_xx = 0.0;  // includes/Module.h:104-110
        eps = (double)FLT_EPSILON;
    }
};
const Consts K;

// ---- ISO4 / ISO8 derivative stencils ---------------------------------------------------
// Each entry lists the "plus" offsets of the central differences of one weight group in
// the exact order the reference sums them; the matching "minus" term is the negated
// offset.  ISO4: src/main_iteration_GPU.cu:765-791 and :914-996 (same per-axis pattern
// for K4 normals and K7 curvature derivatives).
struct Off { int x, y, z; };
const Off ISO4_G2[3][4] = {
    {{1, 1, 0}, {1, -1, 0}, {1, 0, 1}, {1, 0, -1}},   // d/dx
    {{1, 1, 0}, {-1, 1, 0}, {0, 1, 1}, {0, 1, -1}},   // d/dy
    {{1, 0, 1}, {-1, 0, 1}, {0, 1, 1}, {0, -1, 1}}};  // d/dz
const Off ISO4_G1[3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};

// ISO8 (src/Geometry_preprocessing.cpp:232-384): groups of sizes 1,4,4,1,8,12,4.
const int ISO8_N[7] = {1, 4, 4, 1, 8, 12, 4};
const Off ISO8_T[3][34] = {
    {  // d/dx
        {1, 0, 0},
        {1, 1, 0}, {1, -1, 0}, {1, 0, 1}, {1, 0, -1},
        {1, 1, 1}, {1, 1, -1}, {1, -1, 1}, {1, -1, -1},
        {2, 0, 0},
        {2, 1, 0}, {2, -1, 0}, {2, 0, 1}, {2, 0, -1}, {1, 2, 0}, {1, -2, 0}, {1, 0, 2}, {1, 0, -2},
        {2, 1, 1}, {2, 1, -1}, {2, -1, 1}, {2, -1, -1}, {1, 2, 1}, {1, 2, -1}, {1, -2, 1}, {1, -2, -1},
        {1, 1, 2}, {1, 1, -2}, {1, -1, 2}, {1, -1, -2},
        {2, 2, 0}, {2, -2, 0}, {2, 0, 2}, {2, 0, -2}},
    {  // d/dy
        {0, 1, 0},
        {1, 1, 0}, {-1, 1, 0}, {0, 1, 1}, {0, 1, -1},
        {1, 1, 1}, {1, 1, -1}, {-1, 1, -1}, {-1, 1, 1},
        {0, 2, 0},
        {2, 1, 0}, {-2, 1, 0}, {0, 2, 1}, {0, 2, -1}, {1, 2, 0}, {-1, 2, 0}, {0, 1, 2}, {0, 1, -2},
        {2, 1, 1}, {2, 1, -1}, {-2, 1, 1}, {-2, 1, -1}, {1, 2, 1}, {1, 2, -1}, {-1, 2, 1}, {-1, 2, -1},
        {1, 1, 2}, {1, 1, -2}, {-1, 1, 2}, {-1, 1, -2},
        {2, 2, 0}, {-2, 2, 0}, {0, 2, 2}, {0, 2, -2}},
    {  // d/dz
        {0, 0, 1},
        {0, 1, 1}, {0, -1, 1}, {1, 0, 1}, {-1, 0, 1},
        {1, 1, 1}, {1, -1, 1}, {-1, 1, 1}, {-1, -1, 1},
        {0, 0, 2},
        {0, 1, 2}, {0, -1, 2}, {2, 0, 1}, {-2, 0, 1}, {0, 2, 1}, {0, -2, 1}, {1, 0, 2}, {-1, 0, 2},
        {2, 1, 1}, {2, -1, 1}, {-2, 1, 1}, {-2, -1, 1}, {1, 2, 1}, {1, -2, 1}, {-1, 2, 1}, {-1, -2, 1},
        {1, 1, 2}, {1, -1, 2}, {-1, 1, 2}, {-1, -1, 2},
        {0, 2, 2}, {0, -2, 2}, {2, 0, 2}, {-2, 0, 2}}};

// 27-point smoothing stencil order (src/Geometry_preprocessing.cpp:31-34).
const int SX[27] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1};
const int SY[27] = {0, 0, 0, 1, -1, 0, 0, 1, 1, -1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, -1, 1, -1, 1};
const int SZ[27] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, 1, -1, -1, 1, 1, -1, -1, -1, 1, 1, -1, -1, 1, 1, -1};

}  // namespace

// =====================================================================================
//  Simulation state
// =====================================================================================
struct OracleSim {
    oracle_params p;
    int64_t nx, ny, nz;
    int64_t X1, Y1, Z1, X2, Y2, Z2, X4, Y4, Z4;
    int64_t S1, S2, S4;
    std::vector<int32_t> walls;       // s2, 0 fluid / 1 solid
    std::vector<int32_t> walls_type;  // s4, -1 fluid boundary, 0 fluid, 1 solid, 2 solid boundary
    std::vector<double> s_nx, s_ny, s_nz;  // s4 solid normals at fluid-boundary nodes
    std::vector<double> smooth;       // s4: the smoothed solid indicator the solid normals are taken of
    std::vector<double> pdf;          // 38 * s1
    std::vector<double> phi;          // s4
    std::vector<double> cn_x, cn_y, cn_z, c_norm;  // s2
    std::vector<double> curv;         // s1
    std::vector<int32_t> pore_profile_z;
    std::vector<double> w_in;         // X1 * Y1: velocity-inlet profile (plane k = 0 of W_in)
    std::vector<double> conv;         // 39 * X1 * Y1: f_convec, g_convec ([e][y][x]), phi_convec
    int64_t n_solid_boundary = 0, n_fluid_boundary = 0;

    // index helpers: i,j,k are 1-based interior coordinates as in includes/Idx_cpu.h:49-66
    inline int64_t s1(int64_t i, int64_t j, int64_t k) const { return i + X1 * (j + Y1 * k); }
    inline int64_t s2(int64_t i, int64_t j, int64_t k) const { return (i + 1) + X2 * ((j + 1) + Y2 * (k + 1)); }
    inline int64_t s4(int64_t i, int64_t j, int64_t k) const { return (i + 3) + X4 * ((j + 3) + Y4 * (k + 3)); }
    inline int64_t f1(int64_t i, int64_t j, int64_t k, int e, int g) const { return s1(i, j, k) + S1 * (e + 19 * g); }
    inline double& F(int64_t i, int64_t j, int64_t k, int e, int g) { return pdf[f1(i, j, k, e, g)]; }

    bool open_z() const { return p.kper == 0 && p.wall_z_min == 0 && p.wall_z_max == 0; }
    inline int64_t col(int64_t i, int64_t j) const { return i + X1 * j; }
    inline double& CF(int64_t i, int64_t j, int e, int g) { return conv[col(i, j) + X1 * Y1 * (e + 19 * g)]; }
    inline double& CPHI(int64_t i, int64_t j) { return conv[col(i, j) + X1 * Y1 * 38]; }
};

namespace {

// =====================================================================================
//  Geometry: set_walls + geometry_preprocessing_new
//  (src/Misc.cpp:67-91, 195-217; src/Geometry_preprocessing.cpp:29-403)
// =====================================================================================
void preprocess_geometry(OracleSim& S, const int8_t* geo) {
    const int64_t nx = S.nx, ny = S.ny, nz = S.nz;
    std::vector<int32_t> wg(nx * ny * nz);
    for (int64_t n = 0; n < nx * ny * nz; ++n) wg[n] = geo[n] ? 1 : 0;
    auto g0 = [&](int64_t i, int64_t j, int64_t k) -> int32_t& { return wg[(i - 1) + nx * ((j - 1) + ny * (k - 1))]; };
    // domain walls (src/Misc.cpp:67-78)
    for (int64_t k = 1; k <= nz; ++k)
        for (int64_t j = 1; j <= ny; ++j)
            for (int64_t i = 1; i <= nx; ++i) {
                if (S.p.wall_z_min == 1) g0(i, j, 1) = 1;
                if (S.p.wall_z_max == 1) g0(i, j, nz) = 1;
                if (S.p.wall_x_min == 1) g0(1, j, k) = 1;
                if (S.p.wall_x_max == 1) g0(nx, j, k) = 1;
                if (S.p.wall_y_min == 1) g0(i, 1, k) = 1;
                if (S.p.wall_y_max == 1) g0(i, ny, k) = 1;
            }

    // temporary arrays with 10 ghost layers (src/Geometry_preprocessing.cpp:12-20, 37-46)
    const int64_t G = 10;
    const int64_t TX = nx + 2 * G, TY = ny + 2 * G, TZ = nz + 2 * G;
    auto ti = [&](int64_t i, int64_t j, int64_t k) { return (i + G - 1) + TX * ((j + G - 1) + TY * (k + G - 1)); };
    std::vector<int32_t> wt(TX * TY * TZ, 0);
    for (int64_t k = 1; k <= nz; ++k)
        for (int64_t j = 1; j <= ny; ++j)
            for (int64_t i = 1; i <= nx; ++i) wt[ti(i, j, k)] = g0(i, j, k);
    // ghost fill: z, then y, then x (src/Geometry_preprocessing.cpp:56-123)
    for (int64_t j = 1; j <= ny; ++j)
        for (int64_t i = 1; i <= nx; ++i) {
            for (int64_t k = 1 - G; k <= 0; ++k) wt[ti(i, j, k)] = S.p.kper ? wt[ti(i, j, nz + k)] : wt[ti(i, j, 1)];
            for (int64_t k = nz + 1; k <= nz + G; ++k) wt[ti(i, j, k)] = S.p.kper ? wt[ti(i, j, k - nz)] : wt[ti(i, j, nz)];
        }
    for (int64_t k = 1 - G; k <= nz + G; ++k)
        for (int64_t i = 1; i <= nx; ++i) {
            for (int64_t j = 1 - G; j <= 0; ++j) wt[ti(i, j, k)] = S.p.jper ? wt[ti(i, ny + j, k)] : wt[ti(i, 1, k)];
            for (int64_t j = ny + 1; j <= ny + G; ++j) wt[ti(i, j, k)] = S.p.jper ? wt[ti(i, j - ny, k)] : wt[ti(i, ny, k)];
        }
    for (int64_t k = 1 - G; k <= nz + G; ++k)
        for (int64_t j = 1 - G; j <= ny + G; ++j) {
            for (int64_t i = 1 - G; i <= 0; ++i) wt[ti(i, j, k)] = S.p.iper ? wt[ti(nx + i, j, k)] : wt[ti(1, j, k)];
            for (int64_t i = nx + 1; i <= nx + G; ++i) wt[ti(i, j, k)] = S.p.iper ? wt[ti(i - nx, j, k)] : wt[ti(nx, j, k)];
        }
    // walls (s2) keeps the 0/1 flags (src/Geometry_preprocessing.cpp:126-132)
    for (int64_t k = -1; k <= nz + 2; ++k)
        for (int64_t j = -1; j <= ny + 2; ++j)
            for (int64_t i = -1; i <= nx + 2; ++i) S.walls[S.s2(i, j, k)] = wt[ti(i, j, k)];

    // smoothing buffers start from the 0/1 flags (src/Geometry_preprocessing.cpp:135-142)
    std::vector<double> sm1(TX * TY * TZ), sm2(TX * TY * TZ);
    for (int64_t n = 0; n < TX * TY * TZ; ++n) sm1[n] = sm2[n] = (double)wt[n];

    // node classification (src/Geometry_preprocessing.cpp:145-168).  The reference updates
    // in place; 1->2 and 0->-1 keep the tests ">= 1" and "<= 0" true, so a separate output
    // array gives the same result.
    std::vector<int32_t> wc(wt);
#pragma omp parallel for collapse(2) schedule(static)
    for (int64_t k = 2 - G; k <= nz + G - 1; ++k)
        for (int64_t j = 2 - G; j <= ny + G - 1; ++j)
            for (int64_t i = 2 - G; i <= nx + G - 1; ++i) {
                int32_t v = wt[ti(i, j, k)];
                if (v == 1) {
                    for (int q = 1; q <= 18; ++q)
                        if (wt[ti(i + EX[q], j + EY[q], k + EZ[q])] <= 0) { v = 2; break; }
                } else if (v == 0) {
                    for (int q = 1; q <= 18; ++q)
                        if (wt[ti(i + EX[q], j + EY[q], k + EZ[q])] >= 1) { v = -1; break; }
                }
                wc[ti(i, j, k)] = v;
            }
    for (int64_t k = -3; k <= nz + 4; ++k)
        for (int64_t j = -3; j <= ny + 4; ++j)
            for (int64_t i = -3; i <= nx + 4; ++i) S.walls_type[S.s4(i, j, k)] = wc[ti(i, j, k)];

    // 4 Jacobi smoothing sweeps with the D3Q27 filter (src/Geometry_preprocessing.cpp:171-191)
    const double we[4] = {8.0 / 27.0, 2.0 / 27.0, 1.0 / 54.0, 1.0 / 216.0};
    for (int it = 0; it < 4; ++it) {
#pragma omp parallel for collapse(2) schedule(static)
        for (int64_t k = 2 - G; k <= nz + G - 1; ++k)
            for (int64_t j = 2 - G; j <= ny + G - 1; ++j)
                for (int64_t i = 2 - G; i <= nx + G - 1; ++i) {
                    double acc = 0.0;
                    for (int n = 0; n < 27; ++n) {
                        int m = SX[n] * SX[n] + SY[n] * SY[n] + SZ[n] * SZ[n];
                        acc += sm1[ti(i + SX[n], j + SY[n], k + SZ[n])] * we[m];
                    }
                    sm2[ti(i, j, k)] = acc;
                }
#pragma omp parallel for collapse(2) schedule(static)
        for (int64_t k = 2 - G; k <= nz + G - 1; ++k)
            for (int64_t j = 2 - G; j <= ny + G - 1; ++j)
                for (int64_t i = 2 - G; i <= nx + G - 1; ++i) sm1[ti(i, j, k)] = sm2[ti(i, j, k)];
    }

    S.smooth.assign(S.S4, 0.0);
    for (int64_t k = -3; k <= nz + 4; ++k)
        for (int64_t j = -3; j <= ny + 4; ++j)
            for (int64_t i = -3; i <= nx + 4; ++i) S.smooth[S.s4(i, j, k)] = sm2[ti(i, j, k)];

    // solid normals at fluid-boundary nodes (src/Geometry_preprocessing.cpp:207-386)
#pragma omp parallel for collapse(2) schedule(static)
    for (int64_t k = -3; k <= nz + 4; ++k)
        for (int64_t j = -3; j <= ny + 4; ++j)
            for (int64_t i = -3; i <= nx + 4; ++i) {
                if (wc[ti(i, j, k)] != -1) continue;
                double nw[3];
                for (int a = 0; a < 3; ++a) {
                    double tot = 0.0;
                    int t = 0;
                    for (int g = 0; g < 7; ++g) {
                        double acc = 0.0;
                        for (int m = 0; m < ISO8_N[g]; ++m, ++t) {
                            const Off& o = ISO8_T[a][t];
                            double pl = sm2[ti(i + o.x, j + o.y, k + o.z)];
                            double mi = sm2[ti(i - o.x, j - o.y, k - o.z)];
                            acc = (m == 0) ? pl : acc + pl;
                            acc = acc - mi;
                        }
                        double term = K.iso8[g] * acc;
                        tot = (g == 0) ? term : tot + term;
                    }
                    nw[a] = tot;
                }
                double tmp = 1.0 / (std::sqrt(nw[0] * nw[0] + nw[1] * nw[1] + nw[2] * nw[2]) + K.eps);
                S.s_nx[S.s4(i, j, k)] = nw[0] * tmp;
                S.s_ny[S.s4(i, j, k)] = nw[1] * tmp;
                S.s_nz[S.s4(i, j, k)] = nw[2] * tmp;
            }
    S.n_solid_boundary = S.n_fluid_boundary = 0;
    for (int64_t k = -2; k <= nz + 3; ++k)
        for (int64_t j = -2; j <= ny + 3; ++j)
            for (int64_t i = -2; i <= nx + 3; ++i) {
                if (wc[ti(i, j, k)] == 2) S.n_solid_boundary++;
                if (wc[ti(i, j, k)] == -1) S.n_fluid_boundary++;
            }
    // pore profile (src/Misc.cpp:195-217)
    S.pore_profile_z.assign(nz, 0);
    for (int64_t k = 1; k <= nz; ++k)
        for (int64_t j = 1; j <= ny; ++j)
            for (int64_t i = 1; i <= nx; ++i)
                if (S.walls[S.s2(i, j, k)] <= 0) S.pore_profile_z[k - 1] += 1;
}

// =====================================================================================
//  Initial state (src/Init_multiphase.cpp:299-443)
// =====================================================================================
// Option 6's deviate of node n: splitmix64 output function (Steele, Lea and Flood 2014) of
// seed + (n + 1) * 0x9E3779B97F4A7C15, top 53 bits, in [0, 1).  The reference draws
// rand() / RAND_MAX seeded from time(NULL) (src/Init_multiphase.cpp:306, 344-350).
double node_uniform(uint64_t seed, uint64_t n) {
    uint64_t z = seed + (n + 1) * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return (double)(z >> 11) * 0x1p-53;
}

void init_phi(OracleSim& S, int option, double interface_z0, double sa_target = 0.0, uint64_t seed = 0) {
    const int64_t nx = S.nx, ny = S.ny, nz = S.nz;
    std::fill(S.phi.begin(), S.phi.end(), 0.0);
    for (int64_t k = 0; k <= nz + 1; ++k)
        for (int64_t j = 0; j <= ny + 1; ++j)
            for (int64_t i = 0; i <= nx + 1; ++i) {
                double x = (double)i, y = (double)j, z = (double)k;
                double v = 0.0;
                if (option == 1) {
                    v = (z <= interface_z0) ? 1.0 : -1.0;
                } else if (option == 2) {
                    v = (z <= interface_z0) ? -1.0 : 1.0;
                } else if (option == 3 || option == 4 || option == 5) {
                    double cx = (option == 5) ? (double)(nx + 1) * 0.5 : (double)(nx + 1) * 0.0;
                    double dx = x - cx, dz = z - (double)(nz + 1) * 0.5, dy = y - (double)(ny + 1) * 0.5;
                    bool inside = dx * dx + dz * dz + dy * dy <= interface_z0 * interface_z0;
                    if (option == 4) v = inside ? -1.0 : 1.0;
                    else v = inside ? 1.0 : -1.0;
                } else if (option == 6) {
                    const uint64_t n = (uint64_t)(i + (nx + 2) * (j + (ny + 2) * k));  // s1 index
                    v = node_uniform(seed, n) > sa_target ? -1.0 : 1.0;
                }
                S.phi[S.s4(i, j, k)] = v;
            }
    if (S.open_z()) {
        for (int64_t k = -3; k <= nz + 4; ++k)
            for (int64_t j = -3; j <= ny + 4; ++j)
                for (int64_t i = -3; i <= nx + 4; ++i)
                    if ((double)k <= 0) S.phi[S.s4(i, j, k)] = S.p.phi_inlet;
    }
}

// equilibrium populations (src/Init_multiphase.cpp:389-443); rho/u/v/w on s1, null = rest.
// relaxed: the constant terms r * w_q use the collision's own rest weights r0, r1, r2
// (fourth-order moment e2 = 0 instead of the D3Q19 weights' 3 rho); the velocity terms
// are unchanged.  The reference has no such option.
void init_pdf(OracleSim& S, const double* rho, const double* u, const double* v, const double* w,
              bool relaxed = false) {
    const int64_t nx = S.nx, ny = S.ny, nz = S.nz;
#pragma omp parallel for collapse(2) schedule(static)
    for (int64_t k = 0; k <= nz + 1; ++k)
        for (int64_t j = 0; j <= ny + 1; ++j)
            for (int64_t i = 0; i <= nx + 1; ++i) {
                int64_t n = S.s1(i, j, k);
                double uu = u ? u[n] : 0.0, vv = v ? v[n] : 0.0, ww = w ? w[n] : 0.0, rr = rho ? rho[n] : 1.0;
                double usqrt = uu * uu + vv * vv + ww * ww;
                double ph = S.phi[S.s4(i, j, k)];
                double rc[2] = {rr * (1.0 + ph) * 0.5, rr * (1.0 - ph) * 0.5};
                for (int g = 0; g < 2; ++g) {
                    double r = rc[g];
                    S.pdf[n + S.S1 * (0 + 19 * g)] = r * (relaxed ? K.r0 : K.w[0]) + r * K.w[0] * (-1.5 * usqrt);
                    for (int q = 1; q < 19; ++q) {
                        // e.u keeps the reference's operand order: (+-a) (+-b) for diagonals
                        double cu;
                        double comp[3] = {uu, vv, ww};
                        int e[3] = {EX[q], EY[q], EZ[q]};
                        int a = -1, b = -1;
                        for (int d = 0; d < 3; ++d)
                            if (e[d] != 0) { if (a < 0) a = d; else b = d; }
                        double ta = e[a] > 0 ? comp[a] : -comp[a];
                        if (b < 0) cu = ta;
                        else cu = e[b] > 0 ? ta + comp[b] : ta - comp[b];
                        double wq = (q <= 6) ? K.w1 : K.w2;
                        double rq = relaxed ? ((q <= 6) ? K.r1 : K.r2) : wq;
                        S.pdf[n + S.S1 * (q + 19 * g)] = r * rq + r * wq * (3.0 * cu + 4.5 * cu * cu - 1.5 * usqrt);
                    }
                }
            }
}

// =====================================================================================
//  Collision + recolouring of one node (src/main_iteration_GPU.cu:116-345)
//  g1, g2: incoming populations of the two colours by direction;
//  out1, out2: outgoing post-collision populations by direction.
//  Returns phi of the node.
// =====================================================================================
// Relaxation rates of the collision models the reference compiles, selected by its `mrt`
// macro (src/main_iteration_GPU.cu:157-187): 1 bounce-back optimised, 2 original (the
// default), 3 SRT, 4 advection optimised.  s_nu = omega in all four.
struct MrtRates {
    double e, e2, pi, q, t;
};
inline MrtRates mrt_rates(int mrt, double omega) {
    switch (mrt) {
        case 1: {
            const double q = 8.0 * (2.0 - omega) / (8.0 - omega);
            return {omega, omega, omega, q, q};
        }
        case 3: return {omega, omega, omega, omega, omega};
        case 4: return {omega, omega, omega, (6.0 - 3.0 * omega) / (3.0 - omega), omega};
        default: return {1.19, 1.4, 1.4, 1.2, 1.98};
    }
}

inline double collide_node(const oracle_params& P, const double* g1, const double* g2,
                           double cnx, double cny, double cnz, double curv, double cnorm,
                           double* out1, double* out2) {
    double ft[19];
    for (int q = 0; q < 19; ++q) ft[q] = g1[q] + g2[q];
    double rho1 = g1[0], rho2 = g2[0];
    for (int q = 1; q < 19; ++q) { rho1 = rho1 + g1[q]; rho2 = rho2 + g2[q]; }
    const double phi = (rho1 - rho2) / (rho1 + rho2);

    double tmp = 0.5 * P.lbm_gamma * curv * cnorm;
    const double fx = tmp * cnx, fy = tmp * cny, fz = tmp * cnz + P.force_z;

    const double omega = 1.0 / (6.0 / ((1.0 + phi) * P.la_nui1 + (1.0 - phi) * P.la_nui2) + 0.5);
    const MrtRates R = mrt_rates(P.mrt, omega);
    const double s_nu = omega, s_e = R.e, s_e2 = R.e2, s_pi = R.pi, s_q = R.q, s_t = R.t;

    const double den = rho1 + rho2;
    const double ux = ft[1] - ft[2] + ft[7] - ft[8] + ft[9] - ft[10] + ft[11] - ft[12] + ft[13] - ft[14] + 0.5 * fx;
    const double uy = ft[3] - ft[4] + ft[7] + ft[8] - ft[9] - ft[10] + ft[15] - ft[16] + ft[17] - ft[18] + 0.5 * fy;
    const double uz = ft[5] - ft[6] + ft[11] + ft[12] - ft[13] - ft[14] + ft[15] + ft[16] - ft[17] - ft[18] + 0.5 * fz;
    const double u2 = ux * ux + uy * uy + uz * uz;

    double sum1 = ft[1] + ft[2] + ft[3] + ft[4] + ft[5] + ft[6];
    double sum2 = ft[7] + ft[8] + ft[9] + ft[10] + ft[11] + ft[12] + ft[13] + ft[14] + ft[15] + ft[16] + ft[17] + ft[18];
    double sum3 = ft[7] - ft[8] + ft[9] - ft[10] + ft[11] - ft[12] + ft[13] - ft[14];
    double sum4 = ft[7] + ft[8] - ft[9] - ft[10] + ft[15] - ft[16] + ft[17] - ft[18];
    double sum5 = ft[11] + ft[12] - ft[13] - ft[14] + ft[15] + ft[16] - ft[17] - ft[18];
    double sum6 = 2.0 * (ft[1] + ft[2]) - ft[3] - ft[4] - ft[5] - ft[6];
    double sum7 = ft[7] + ft[8] + ft[9] + ft[10] + ft[11] + ft[12] + ft[13] + ft[14] - 2.0 * (ft[15] + ft[16] + ft[17] + ft[18]);
    double sum8 = ft[3] + ft[4] - ft[5] - ft[6];
    double sum9 = ft[7] + ft[8] + ft[9] + ft[10] - ft[11] - ft[12] - ft[13] - ft[14];

    double m_rho = den;
    double m_e = -30.0 * ft[0] - 11.0 * sum1 + 8.0 * sum2;
    double m_e2 = 12.0 * ft[0] - 4.0 * sum1 + sum2;
    double m_jx = ft[1] - ft[2] + sum3;
    double m_qx = -4.0 * (ft[1] - ft[2]) + sum3;
    double m_jy = ft[3] - ft[4] + sum4;
    double m_qy = -4.0 * (ft[3] - ft[4]) + sum4;
    double m_jz = ft[5] - ft[6] + sum5;
    double m_qz = -4.0 * (ft[5] - ft[6]) + sum5;
    double m_3pxx = sum6 + sum7;
    double m_3pixx = -2.0 * sum6 + sum7;
    double m_pww = sum8 + sum9;
    double m_piww = -2.0 * sum8 + sum9;
    double m_pxy = ft[7] - ft[8] - ft[9] + ft[10];
    double m_pyz = ft[15] - ft[16] - ft[17] + ft[18];
    double m_pzx = ft[11] - ft[12] - ft[13] + ft[14];
    double m_tx = ft[7] - ft[8] + ft[9] - ft[10] - ft[11] + ft[12] - ft[13] + ft[14];
    double m_ty = -ft[7] - ft[8] + ft[9] + ft[10] + ft[15] - ft[16] + ft[17] - ft[18];
    double m_tz = ft[11] + ft[12] - ft[13] - ft[14] - ft[15] - ft[16] + ft[17] + ft[18];

    // relaxation with forcing (src/main_iteration_GPU.cu:228-246)
    const double c23 = -0.666666666666666667, c13 = 0.333333333333333333;
    m_e = m_e - s_e * (m_e - (-11.0 * den + 19.0 * u2)) + (38.0 - 19.0 * s_e) * (fx * ux + fy * uy + fz * uz);
    m_e2 = m_e2 - s_e2 * (m_e2 - (K.e2c1 * den + K.e2c2 * u2)) + (-11.0 + 5.5 * s_e2) * (fx * ux + fy * uy + fz * uz);
    m_jx = m_jx + fx;
    m_qx = m_qx - s_q * (m_qx - (c23 * ux)) + (c23 + c13 * s_q) * fx;
    m_jy = m_jy + fy;
    m_qy = m_qy - s_q * (m_qy - (c23 * uy)) + (c23 + c13 * s_q) * fy;
    m_jz = m_jz + fz;
    m_qz = m_qz - s_q * (m_qz - (c23 * uz)) + (c23 + c13 * s_q) * fz;
    m_3pxx = m_3pxx - s_nu * (m_3pxx - (3.0 * ux * ux - u2)) + (2.0 - s_nu) * (2.0 * fx * ux - fy * uy - fz * uz);
    m_3pixx = m_3pixx - s_pi * (m_3pixx - K.omega_xx * (3.0 * ux * ux - u2)) + (1.0 - 0.5 * s_pi) * (-2.0 * fx * ux + fy * uy + fz * uz);
    m_pww = m_pww - s_nu * (m_pww - (uy * uy - uz * uz)) + (2.0 - s_nu) * (fy * uy - fz * uz);
    m_piww = m_piww - s_pi * (m_piww - K.omega_xx * (uy * uy - uz * uz)) + (1.0 - 0.5 * s_pi) * (-fy * uy + fz * uz);
    m_pxy = m_pxy - s_nu * (m_pxy - (ux * uy)) + (1.0 - 0.5 * s_nu) * (fx * uy + fy * ux);
    m_pyz = m_pyz - s_nu * (m_pyz - (uy * uz)) + (1.0 - 0.5 * s_nu) * (fy * uz + fz * uy);
    m_pzx = m_pzx - s_nu * (m_pzx - (ux * uz)) + (1.0 - 0.5 * s_nu) * (fx * uz + fz * ux);
    m_tx = m_tx - s_t * (m_tx);
    m_ty = m_ty - s_t * (m_ty);
    m_tz = m_tz - s_t * (m_tz);

    // back to populations (src/main_iteration_GPU.cu:250-297)
    m_rho = K.c19 * m_rho;
    m_e = K.c2394 * m_e;
    m_e2 = K.c252 * m_e2;
    m_jx = 0.1 * m_jx;  m_qx = 0.025 * m_qx;
    m_jy = 0.1 * m_jy;  m_qy = 0.025 * m_qy;
    m_jz = 0.1 * m_jz;  m_qz = 0.025 * m_qz;
    m_3pxx = 2.0 * K.c72 * m_3pxx;
    m_3pixx = K.c72 * m_3pixx;
    m_pww = 6.0 * K.c72 * m_pww;
    m_piww = 3.0 * K.c72 * m_piww;
    m_pxy = 0.25 * m_pxy;  m_pyz = 0.25 * m_pyz;  m_pzx = 0.25 * m_pzx;
    m_tx = 0.125 * m_tx;   m_ty = 0.125 * m_ty;   m_tz = 0.125 * m_tz;
    sum1 = m_rho - 11.0 * m_e - 4.0 * m_e2;
    sum2 = 2.0 * m_3pxx - 4.0 * m_3pixx;
    sum3 = m_pww - 2.0 * m_piww;
    sum4 = m_rho + 8.0 * m_e + m_e2;
    sum5 = m_jx + m_qx;
    sum6 = m_jy + m_qy;
    sum7 = m_jz + m_qz;
    sum8 = m_3pxx + m_3pixx;
    sum9 = m_pww + m_piww;
    double f[19];
    f[0] = m_rho - 30.0 * m_e + 12.0 * m_e2;
    f[1] = sum1 + m_jx - 4.0 * m_qx + sum2;
    f[2] = sum1 - m_jx + 4.0 * m_qx + sum2;
    f[3] = sum1 + m_jy - 4.0 * m_qy - 0.5 * sum2 + sum3;
    f[4] = sum1 - m_jy + 4.0 * m_qy - 0.5 * sum2 + sum3;
    f[5] = sum1 + m_jz - 4.0 * m_qz - 0.5 * sum2 - sum3;
    f[6] = sum1 - m_jz + 4.0 * m_qz - 0.5 * sum2 - sum3;
    f[7] = sum4 + sum5 + sum6 + sum8 + sum9 + m_pxy + m_tx - m_ty;
    f[8] = sum4 - sum5 + sum6 + sum8 + sum9 - m_pxy - m_tx - m_ty;
    f[9] = sum4 + sum5 - sum6 + sum8 + sum9 - m_pxy + m_tx + m_ty;
    f[10] = sum4 - sum5 - sum6 + sum8 + sum9 + m_pxy - m_tx + m_ty;
    f[11] = sum4 + sum5 + sum7 + sum8 - sum9 + m_pzx - m_tx + m_tz;
    f[12] = sum4 - sum5 + sum7 + sum8 - sum9 - m_pzx + m_tx + m_tz;
    f[13] = sum4 + sum5 - sum7 + sum8 - sum9 - m_pzx - m_tx - m_tz;
    f[14] = sum4 - sum5 - sum7 + sum8 - sum9 + m_pzx + m_tx - m_tz;
    f[15] = sum4 + sum6 + sum7 - sum8 * 2.0 + m_pyz + m_ty - m_tz;
    f[16] = sum4 - sum6 + sum7 - sum8 * 2.0 - m_pyz - m_ty - m_tz;
    f[17] = sum4 + sum6 - sum7 - sum8 * 2.0 - m_pyz + m_ty + m_tz;
    f[18] = sum4 - sum6 - sum7 - sum8 * 2.0 + m_pyz - m_ty + m_tz;

    // Latva-Kokko--Rothman recolouring (src/main_iteration_GPU.cu:302-345)
    const double tmp1 = rho1 / den;
    out1[0] = tmp1 * f[0];
    out2[0] = f[0] * (1.0 - tmp1);
    tmp = rho1 * rho2 * P.lbm_beta / den;
    const double n3[3] = {cnx, cny, cnz};
    for (int q = 1; q < 19; ++q) {
        // e_q . n with the reference's operand order: +-n_a then +-n_b
        const int e[3] = {EX[q], EY[q], EZ[q]};
        int a = -1, b = -1;
        for (int d = 0; d < 3; ++d)
            if (e[d] != 0) { if (a < 0) a = d; else b = d; }
        double proj = e[a] > 0 ? n3[a] : -n3[a];
        if (b >= 0) proj = e[b] > 0 ? proj + n3[b] : proj - n3[b];
        const double wq = (b < 0) ? K.w1 : K.rk2;
        out1[q] = tmp1 * f[q] + wq * tmp * (proj);
        out2[q] = f[q] - out1[q];
    }
    return phi;
}

// =====================================================================================
//  AA-pattern collide/stream sweeps (src/main_iteration_GPU.cu:56-388 odd, 395-726 even)
// =====================================================================================
void collide_sweep(OracleSim& S, bool odd) {
    const int64_t nx = S.nx, ny = S.ny, nz = S.nz;
#pragma omp parallel for collapse(2) schedule(static)
    for (int64_t k = 1; k <= nz; ++k)
        for (int64_t j = 1; j <= ny; ++j)
            for (int64_t i = 1; i <= nx; ++i) {
                if (S.walls[S.s2(i, j, k)] != 0) continue;
                double g1[19], g2[19], o1[19], o2[19];
                for (int q = 0; q < 19; ++q) {
                    if (odd) {  // pull (x - e_q).slot q
                        int64_t n = S.s1(i - EX[q], j - EY[q], k - EZ[q]);
                        g1[q] = S.pdf[n + S.S1 * q];
                        g2[q] = S.pdf[n + S.S1 * (q + 19)];
                    } else {    // own node, slot opp(q)
                        int64_t n = S.s1(i, j, k);
                        g1[q] = S.pdf[n + S.S1 * OPP[q]];
                        g2[q] = S.pdf[n + S.S1 * (OPP[q] + 19)];
                    }
                }
                const int64_t n2 = S.s2(i, j, k);
                double ph = collide_node(S.p, g1, g2, S.cn_x[n2], S.cn_y[n2], S.cn_z[n2],
                                         S.curv[S.s1(i, j, k)], S.c_norm[n2], o1, o2);
                S.phi[S.s4(i, j, k)] = ph;
                for (int q = 0; q < 19; ++q) {
                    if (odd) {  // push to (x + e_q).slot opp(q)
                        int64_t n = S.s1(i + EX[q], j + EY[q], k + EZ[q]);
                        S.pdf[n + S.S1 * OPP[q]] = o1[q];
                        S.pdf[n + S.S1 * (OPP[q] + 19)] = o2[q];
                    } else {
                        int64_t n = S.s1(i, j, k);
                        S.pdf[n + S.S1 * q] = o1[q];
                        S.pdf[n + S.S1 * (q + 19)] = o2[q];
                    }
                }
            }
}

// =====================================================================================
//  Periodic ghost copies (src/main_iteration_GPU.cu:1529-1733)
// =====================================================================================
void periodic_z(OracleSim& S, bool odd) {
    const int64_t nx = S.nx, ny = S.ny, nz = S.nz;
    const int up[5] = {6, 14, 13, 18, 17}, dn[5] = {5, 11, 12, 15, 16};
    for (int64_t j = 1; j <= ny; ++j)
        for (int64_t i = 1; i <= nx; ++i)
            for (int g = 0; g < 2; ++g)
                for (int m = 0; m < 5; ++m) {
                    if (!odd) {
                        S.F(i, j, nz + 1, up[m], g) = S.F(i, j, 1, up[m], g);
                        S.F(i, j, 0, dn[m], g) = S.F(i, j, nz, dn[m], g);
                    } else {
                        S.F(i, j, 1, up[m], g) = S.F(i, j, nz + 1, up[m], g);
                        S.F(i, j, nz, dn[m], g) = S.F(i, j, 0, dn[m], g);
                    }
                }
}
void periodic_y(OracleSim& S, bool odd) {
    const int64_t nx = S.nx, ny = S.ny, nz = S.nz;
    const int up[5] = {4, 10, 9, 16, 18}, dn[5] = {3, 7, 8, 15, 17};
    for (int64_t k = 1; k <= nz; ++k)
        for (int64_t i = 1; i <= nx; ++i)
            for (int g = 0; g < 2; ++g)
                for (int m = 0; m < 5; ++m) {
                    if (!odd) {
                        S.F(i, ny + 1, k, up[m], g) = S.F(i, 1, k, up[m], g);
                        S.F(i, 0, k, dn[m], g) = S.F(i, ny, k, dn[m], g);
                    } else {
                        S.F(i, 1, k, up[m], g) = S.F(i, ny + 1, k, up[m], g);
                        S.F(i, ny, k, dn[m], g) = S.F(i, 0, k, dn[m], g);
                    }
                }
}
void periodic_zy(OracleSim& S, bool odd) {
    const int64_t nx = S.nx, ny = S.ny, nz = S.nz;
    for (int64_t i = 1; i <= nx; ++i)
        for (int g = 0; g < 2; ++g) {
            if (!odd) {
                S.F(i, ny + 1, nz + 1, 18, g) = S.F(i, 1, 1, 18, g);
                S.F(i, ny + 1, 0, 16, g) = S.F(i, 1, nz, 16, g);
                S.F(i, 0, nz + 1, 17, g) = S.F(i, ny, 1, 17, g);
                S.F(i, 0, 0, 15, g) = S.F(i, ny, nz, 15, g);
            } else {
                S.F(i, 1, 1, 18, g) = S.F(i, ny + 1, nz + 1, 18, g);
                S.F(i, 1, nz, 16, g) = S.F(i, ny + 1, 0, 16, g);
                S.F(i, ny, 1, 17, g) = S.F(i, 0, nz + 1, 17, g);
                S.F(i, ny, nz, 15, g) = S.F(i, 0, 0, 15, g);
            }
        }
}
// faces (z, y) or, with edges = true, only the y-z edge blocks
void periodic_phi(OracleSim& S, bool z, bool y, bool edges = false) {
    const int64_t nx = S.nx, ny = S.ny, nz = S.nz;
    const int ov = 4;
    auto P = [&](int64_t i, int64_t j, int64_t k) -> double& { return S.phi[S.s4(i, j, k)]; };
    if (z && !edges)
        for (int64_t j = 1; j <= ny; ++j)
            for (int64_t i = 1; i <= nx; ++i)
                for (int k = 1; k <= ov; ++k) {
                    P(i, j, k + nz) = P(i, j, k);
                    P(i, j, k - ov) = P(i, j, nz + k - ov);
                }
    if (y && !edges)
        for (int64_t k = 1; k <= nz; ++k)
            for (int64_t i = 1; i <= nx; ++i)
                for (int j = 1; j <= ov; ++j) {
                    P(i, j + ny, k) = P(i, j, k);
                    P(i, j - ov, k) = P(i, ny + j - ov, k);
                }
    if (edges)
        for (int64_t i = 1; i <= nx; ++i)
            for (int k = 1; k <= ov; ++k)
                for (int j = 1; j <= ov; ++j) {
                    P(i, j - ov, k - ov) = P(i, ny + j - ov, nz + k - ov);
                    P(i, j + ny, k - ov) = P(i, j, nz + k - ov);
                    P(i, j + ny, k + nz) = P(i, j, k);
                    P(i, j - ov, k + nz) = P(i, ny + j - ov, k);
                }
}

// =====================================================================================
//  Zou-He pressure inlet (k = 1) and outlet (k = nz)
//  (src/main_iteration_GPU.cu:1085-1241 inlet, 1363-1524 outlet)
//  The sums are written slot by slot: their order differs between the before-odd and the
//  after-odd variants, and the rounding must follow each.
// =====================================================================================
void inlet_zou_he(OracleSim& S, bool after_odd) {
    const int64_t nx = S.nx, ny = S.ny;
    const double c3 = 0.333333333333333333, c6 = 0.166666666666666667;
    for (int64_t j = 1; j <= ny; ++j)
        for (int64_t i = 1; i <= nx; ++i) {
            const int w = S.walls[S.s2(i, j, 1)];
            double& p0 = S.phi[S.s4(i, j, 0)];
            p0 = S.p.phi_inlet * (1 - w) + p0 * w;
            S.phi[S.s4(i, j, -1)] = p0;
            S.phi[S.s4(i, j, -2)] = p0;
            S.phi[S.s4(i, j, -3)] = p0;
            double rhoT[2];
            rhoT[1] = S.p.rho_in;
            rhoT[0] = S.p.rho_in * S.p.sa_inject;
            rhoT[1] = rhoT[1] - rhoT[0];
            for (int g = 0; g < 2; ++g) {
                auto f = [&](int64_t a, int64_t b, int64_t c, int e) -> double& { return S.F(a, b, c, e, g); };
                if (!after_odd) {
                    double t = (rhoT[g] - (f(i, j, 1, 0) + f(i - 1, j, 1, 1) + f(i + 1, j, 1, 2) + f(i, j - 1, 1, 3) +
                                           f(i, j + 1, 1, 4) + f(i - 1, j - 1, 1, 7) + f(i + 1, j - 1, 1, 8) +
                                           f(i - 1, j + 1, 1, 9) + f(i + 1, j + 1, 1, 10) +
                                           2.0 * (f(i, j, 2, 6) + f(i + 1, j, 2, 14) + f(i - 1, j, 2, 13) +
                                                  f(i, j + 1, 2, 18) + f(i, j - 1, 2, 17)))) * S.p.relaxation;
                    double tnx = 0.5 * (f(i - 1, j, 1, 1) + f(i - 1, j - 1, 1, 7) + f(i - 1, j + 1, 1, 9) -
                                        (f(i + 1, j, 1, 2) + f(i + 1, j - 1, 1, 8) + f(i + 1, j + 1, 1, 10)));
                    double tny = 0.5 * (f(i, j - 1, 1, 3) + f(i - 1, j - 1, 1, 7) + f(i + 1, j - 1, 1, 8) -
                                        (f(i, j + 1, 1, 4) + f(i + 1, j + 1, 1, 10) + f(i - 1, j + 1, 1, 9)));
                    f(i, j, 0, 5) = (f(i, j, 2, 6) + c3 * t) * (1 - w) + f(i, j, 0, 5) * w;
                    f(i - 1, j, 0, 11) = (f(i + 1, j, 2, 14) + c6 * t - tnx) * (1 - w) + f(i - 1, j, 0, 11) * w;
                    f(i + 1, j, 0, 12) = (f(i - 1, j, 2, 13) + c6 * t + tnx) * (1 - w) + f(i + 1, j, 0, 12) * w;
                    f(i, j - 1, 0, 15) = (f(i, j + 1, 2, 18) + c6 * t - tny) * (1 - w) + f(i, j - 1, 0, 15) * w;
                    f(i, j + 1, 0, 16) = (f(i, j - 1, 2, 17) + c6 * t + tny) * (1 - w) + f(i, j + 1, 0, 16) * w;
                } else {
                    double t = (rhoT[g] - (f(i, j, 1, 0) + f(i, j, 1, 2) + f(i, j, 1, 1) + f(i, j, 1, 4) +
                                           f(i, j, 1, 3) + f(i, j, 1, 8) + f(i, j, 1, 7) + f(i, j, 1, 10) +
                                           f(i, j, 1, 9) +
                                           2.0 * (f(i, j, 1, 5) + f(i, j, 1, 11) + f(i, j, 1, 12) +
                                                  f(i, j, 1, 15) + f(i, j, 1, 16)))) * S.p.relaxation;
                    double tnx = 0.5 * (f(i, j, 1, 2) + f(i, j, 1, 8) + f(i, j, 1, 10) -
                                        (f(i, j, 1, 1) + f(i, j, 1, 7) + f(i, j, 1, 9)));
                    double tny = 0.5 * (f(i, j, 1, 4) + f(i, j, 1, 9) + f(i, j, 1, 10) -
                                        (f(i, j, 1, 3) + f(i, j, 1, 8) + f(i, j, 1, 7)));
                    f(i, j, 1, 6) = (f(i, j, 1, 5) + c3 * t) * (1 - w) + f(i, j, 1, 6) * w;
                    f(i, j, 1, 13) = (f(i, j, 1, 12) + c6 * t + tnx) * (1 - w) + f(i, j, 1, 13) * w;
                    f(i, j, 1, 14) = (f(i, j, 1, 11) + c6 * t - tnx) * (1 - w) + f(i, j, 1, 14) * w;
                    f(i, j, 1, 17) = (f(i, j, 1, 16) + c6 * t + tny) * (1 - w) + f(i, j, 1, 17) * w;
                    f(i, j, 1, 18) = (f(i, j, 1, 15) + c6 * t - tny) * (1 - w) + f(i, j, 1, 18) * w;
                }
            }
        }
}

void outlet_zou_he(OracleSim& S, bool after_odd) {
    const int64_t nx = S.nx, ny = S.ny, nz = S.nz;
    const double c3 = 0.333333333333333333, c6 = 0.166666666666666667;
    for (int64_t j = 1; j <= ny; ++j)
        for (int64_t i = 1; i <= nx; ++i) {
            const int w = S.walls[S.s2(i, j, nz)];
            const double ph = S.phi[S.s4(i, j, nz)];
            for (int m = 1; m <= 4; ++m) S.phi[S.s4(i, j, nz + m)] = ph;
            auto f = [&](int64_t a, int64_t b, int64_t c, int e, int g) -> double& { return S.F(a, b, c, e, g); };
            double t1;
            if (!after_odd) {
                t1 = f(i, j, nz, 0, 0) + f(i - 1, j, nz, 1, 0) + f(i + 1, j, nz, 2, 0) + f(i, j - 1, nz, 3, 0) +
                     f(i, j + 1, nz, 4, 0) + f(i - 1, j - 1, nz, 7, 0) + f(i + 1, j - 1, nz, 8, 0) +
                     f(i - 1, j + 1, nz, 9, 0) + f(i + 1, j + 1, nz, 10, 0) +
                     2.0 * (f(i, j, nz - 1, 5, 0) + f(i - 1, j, nz - 1, 11, 0) + f(i + 1, j, nz - 1, 12, 0) +
                            f(i, j - 1, nz - 1, 15, 0) + f(i, j + 1, nz - 1, 16, 0)) +
                     f(i, j, nz, 0, 1) + f(i - 1, j, nz, 1, 1) + f(i + 1, j, nz, 2, 1) + f(i, j - 1, nz, 3, 1) +
                     f(i, j + 1, nz, 4, 1) + f(i - 1, j - 1, nz, 7, 1) + f(i + 1, j - 1, nz, 8, 1) +
                     f(i - 1, j + 1, nz, 9, 1) + f(i + 1, j + 1, nz, 10, 1) +
                     2.0 * (f(i, j, nz - 1, 5, 1) + f(i - 1, j, nz - 1, 11, 1) + f(i + 1, j, nz - 1, 12, 1) +
                            f(i, j - 1, nz - 1, 15, 1) + f(i, j + 1, nz - 1, 16, 1)) -
                     S.p.rho_out;
            } else {
                t1 = f(i, j, nz, 0, 0) + f(i, j, nz, 2, 0) + f(i, j, nz, 1, 0) + f(i, j, nz, 4, 0) +
                     f(i, j, nz, 3, 0) + f(i, j, nz, 8, 0) + f(i, j, nz, 7, 0) + f(i, j, nz, 10, 0) +
                     f(i, j, nz, 9, 0) +
                     2.0 * (f(i, j, nz, 6, 0) + f(i, j, nz, 14, 0) + f(i, j, nz, 13, 0) + f(i, j, nz, 18, 0) +
                            f(i, j, nz, 17, 0)) +
                     f(i, j, nz, 0, 1) + f(i, j, nz, 2, 1) + f(i, j, nz, 1, 1) + f(i, j, nz, 4, 1) +
                     f(i, j, nz, 3, 1) + f(i, j, nz, 8, 1) + f(i, j, nz, 7, 1) + f(i, j, nz, 10, 1) +
                     f(i, j, nz, 9, 1) +
                     2.0 * (f(i, j, nz, 6, 1) + f(i, j, nz, 14, 1) + f(i, j, nz, 13, 1) + f(i, j, nz, 18, 1) +
                            f(i, j, nz, 17, 1)) -
                     S.p.rho_out;
            }
            double t2 = t1 * 0.5 * (1.0 - ph);   // fluid 2 net flux
            t1 = t1 - t2;                        // fluid 1 net flux
            const double tt[2] = {t1, t2};
            for (int g = 0; g < 2; ++g) {
                const double t = tt[g];
                if (!after_odd) {
                    double tnx = 0.5 * (f(i - 1, j, nz, 1, g) + f(i - 1, j - 1, nz, 7, g) + f(i - 1, j + 1, nz, 9, g) -
                                        (f(i + 1, j, nz, 2, g) + f(i + 1, j - 1, nz, 8, g) + f(i + 1, j + 1, nz, 10, g)));
                    double tny = 0.5 * (f(i, j - 1, nz, 3, g) + f(i - 1, j - 1, nz, 7, g) + f(i + 1, j - 1, nz, 8, g) -
                                        (f(i, j + 1, nz, 4, g) + f(i + 1, j + 1, nz, 10, g) + f(i - 1, j + 1, nz, 9, g)));
                    f(i, j, nz + 1, 6, g) = (f(i, j, nz - 1, 5, g) - c3 * t) * (1 - w) + f(i, j, nz + 1, 6, g) * w;
                    f(i - 1, j, nz + 1, 13, g) = (f(i + 1, j, nz - 1, 12, g) - c6 * t - tnx) * (1 - w) + f(i - 1, j, nz + 1, 13, g) * w;
                    f(i + 1, j, nz + 1, 14, g) = (f(i - 1, j, nz - 1, 11, g) - c6 * t + tnx) * (1 - w) + f(i + 1, j, nz + 1, 14, g) * w;
                    f(i, j - 1, nz + 1, 17, g) = (f(i, j + 1, nz - 1, 16, g) - c6 * t - tny) * (1 - w) + f(i, j - 1, nz + 1, 17, g) * w;
                    f(i, j + 1, nz + 1, 18, g) = (f(i, j - 1, nz - 1, 15, g) - c6 * t + tny) * (1 - w) + f(i, j + 1, nz + 1, 18, g) * w;
                } else {
                    double tnx = 0.5 * (f(i, j, nz, 2, g) + f(i, j, nz, 8, g) + f(i, j, nz, 10, g) -
                                        (f(i, j, nz, 1, g) + f(i, j, nz, 7, g) + f(i, j, nz, 9, g)));
                    double tny = 0.5 * (f(i, j, nz, 4, g) + f(i, j, nz, 10, g) + f(i, j, nz, 9, g) -
                                        (f(i, j, nz, 3, g) + f(i, j, nz, 7, g) + f(i, j, nz, 8, g)));
                    f(i, j, nz, 5, g) = (f(i, j, nz, 6, g) - c3 * t) * (1 - w) + f(i, j, nz, 5, g) * w;
                    f(i, j, nz, 11, g) = (f(i, j, nz, 14, g) - c6 * t + tnx) * (1 - w) + f(i, j, nz, 11, g) * w;
                    f(i, j, nz, 12, g) = (f(i, j, nz, 13, g) - c6 * t - tnx) * (1 - w) + f(i, j, nz, 12, g) * w;
                    f(i, j, nz, 15, g) = (f(i, j, nz, 18, g) - c6 * t + tny) * (1 - w) + f(i, j, nz, 15, g) * w;
                    f(i, j, nz, 16, g) = (f(i, j, nz, 17, g) - c6 * t - tny) * (1 - w) + f(i, j, nz, 16, g) * w;
                }
            }
        }
}

// =====================================================================================
//  Colour-gradient chain (src/Phase_gradient.cpp:15-291, GPU twin
//  src/main_iteration_GPU.cu:732-1003).  Ranges: K3 [-2,n+3], K4/K5 [-1,n+2],
//  K6 [0,n+1], K7 [1,n].
// =====================================================================================
inline double iso4_deriv(const std::vector<double>& a, const OracleSim& S, int axis,
                         int64_t i, int64_t j, int64_t k, bool s4idx) {
    auto at = [&](int64_t x, int64_t y, int64_t z) { return s4idx ? a[S.s4(x, y, z)] : a[S.s2(x, y, z)]; };
    const Off g1 = ISO4_G1[axis];
    double d1 = at(i + g1.x, j + g1.y, k + g1.z) - at(i - g1.x, j - g1.y, k - g1.z);
    double acc = 0.0;
    for (int m = 0; m < 4; ++m) {
        const Off o = ISO4_G2[axis][m];
        double pl = at(i + o.x, j + o.y, k + o.z), mi = at(i - o.x, j - o.y, k - o.z);
        acc = (m == 0) ? pl : acc + pl;
        acc = acc - mi;
    }
    return K.iso4[0] * d1 + K.iso4[1] * acc;
}

void gradient_chain(OracleSim& S) {
    const int64_t nx = S.nx, ny = S.ny, nz = S.nz;
    // K3: phi at solid-boundary nodes
#pragma omp parallel for collapse(2) schedule(static)
    for (int64_t k = -2; k <= nz + 3; ++k)
        for (int64_t j = -2; j <= ny + 3; ++j)
            for (int64_t i = -2; i <= nx + 3; ++i) {
                if (S.walls_type[S.s4(i, j, k)] != 2) continue;
                double ps = 0.0, ws = 0.0;
                for (int q = 1; q < 19; ++q) {
                    int64_t n = S.s4(i + EX[q], j + EY[q], k + EZ[q]);
                    if (S.walls_type[n] <= 0) { ps += S.phi[n] * K.w[q]; ws += K.w[q]; }
                }
                S.phi[S.s4(i, j, k)] = ps / ws;
            }
    // K4: normals; K5 is applied separately below, as in the reference
#pragma omp parallel for collapse(2) schedule(static)
    for (int64_t k = -1; k <= nz + 2; ++k)
        for (int64_t j = -1; j <= ny + 2; ++j)
            for (int64_t i = -1; i <= nx + 2; ++i) {
                const int64_t n = S.s2(i, j, k);
                double cx = iso4_deriv(S.phi, S, 0, i, j, k, true);
                double cy = iso4_deriv(S.phi, S, 1, i, j, k, true);
                double cz = iso4_deriv(S.phi, S, 2, i, j, k, true);
                double cn = std::sqrt(cx * cx + cy * cy + cz * cz);
                if (cn < 1e-6 || S.walls[n] == 1) {
                    S.cn_x[n] = 0.0; S.cn_y[n] = 0.0; S.cn_z[n] = 0.0; S.c_norm[n] = 0.0;
                } else {
                    S.cn_x[n] = cx / cn; S.cn_y[n] = cy / cn; S.cn_z[n] = cz / cn; S.c_norm[n] = cn;
                }
            }
    // K5: geometrical wetting (src/Phase_gradient.cpp:223-291)
    const double lambda = 0.5, leps = 1e-6, ct = S.p.cos_theta;
#pragma omp parallel for collapse(2) schedule(static)
    for (int64_t k = -1; k <= nz + 2; ++k)
        for (int64_t j = -1; j <= ny + 2; ++j)
            for (int64_t i = -1; i <= nx + 2; ++i) {
                const int64_t n4 = S.s4(i, j, k), n2 = S.s2(i, j, k);
                if (S.walls_type[n4] != -1 || !(S.c_norm[n2] > leps)) continue;
                const double nwx = S.s_nx[n4], nwy = S.s_ny[n4], nwz = S.s_nz[n4];
                double vx0 = S.cn_x[n2], vy0 = S.cn_y[n2], vz0 = S.cn_z[n2];
                double vx1 = vx0 - lambda * (vx0 + nwx);
                double vy1 = vy0 - lambda * (vy0 + nwy);
                double vz1 = vz0 - lambda * (vz0 + nwz);
                double e0 = (nwx * vx0 + nwy * vy0 + nwz * vz0) - ct;
                if (!((std::fabs(vx0 + nwx) + std::fabs(vy0 + nwy) + std::fabs(vz0 + nwz) > leps ||
                       std::fabs(vx0 - nwx) + std::fabs(vy0 - nwy) + std::fabs(vz0 - nwz) > leps) && e0 > leps))
                    continue;
                double e1 = (nwx * vx1 + nwy * vy1 + nwz * vz1) - std::sqrt(vx1 * vx1 + vy1 * vy1 + vz1 * vz1) * ct;
                double t = 1.0 / (e1 - e0);
                double vx2 = t * (vx0 * e1 - vx1 * e0);
                double vy2 = t * (vy0 * e1 - vy1 * e0);
                double vz2 = t * (vz0 * e1 - vz1 * e0);
                double e2 = (nwx * vx2 + nwy * vy2 + nwz * vz2) - std::sqrt(vx2 * vx2 + vy2 * vy2 + vz2 * vz2) * ct;
                if (e2 > leps) {
                    for (int it = 2; it <= 4; ++it) {
                        vx0 = vx1; vy0 = vy1; vz0 = vz1;
                        vx1 = vx2; vy1 = vy2; vz1 = vz2;
                        e0 = (nwx * vx0 + nwy * vy0 + nwz * vz0) - std::sqrt(vx0 * vx0 + vy0 * vy0 + vz0 * vz0) * ct;
                        e1 = (nwx * vx1 + nwy * vy1 + nwz * vz1) - std::sqrt(vx1 * vx1 + vy1 * vy1 + vz1 * vz1) * ct;
                        t = 1.0 / (e1 - e0);
                        if (std::isinf(t)) break;
                        vx2 = t * (vx0 * e1 - vx1 * e0);
                        vy2 = t * (vy0 * e1 - vy1 * e0);
                        vz2 = t * (vz0 * e1 - vz1 * e0);
                        e2 = (nwx * vx2 + nwy * vy2 + nwz * vz2) - std::sqrt(vx2 * vx2 + vy2 * vy2 + vz2 * vz2) * ct;
                    }
                }
                t = 1.0 / (1e-30 + std::sqrt(vx2 * vx2 + vy2 * vy2 + vz2 * vz2));
                S.cn_x[n2] = vx2 * t; S.cn_y[n2] = vy2 * t; S.cn_z[n2] = vz2 * t;
            }
    // K6: normals at solid-boundary nodes
#pragma omp parallel for collapse(2) schedule(static)
    for (int64_t k = 0; k <= nz + 1; ++k)
        for (int64_t j = 0; j <= ny + 1; ++j)
            for (int64_t i = 0; i <= nx + 1; ++i) {
                if (S.walls_type[S.s4(i, j, k)] != 2) continue;
                double sx = 0.0, sy = 0.0, sz = 0.0, ws = 0.0;
                for (int q = 1; q < 19; ++q) {
                    int64_t a = i + EX[q], b = j + EY[q], c = k + EZ[q];
                    if (S.walls_type[S.s4(a, b, c)] <= 0) {
                        int64_t n = S.s2(a, b, c);
                        sx += S.cn_x[n] * K.w[q]; sy += S.cn_y[n] * K.w[q]; sz += S.cn_z[n] * K.w[q];
                        ws += K.w[q];
                    }
                }
                int64_t n = S.s2(i, j, k);
                S.cn_x[n] = sx / ws; S.cn_y[n] = sy / ws; S.cn_z[n] = sz / ws;
            }
    // K7: curvature
#pragma omp parallel for collapse(2) schedule(static)
    for (int64_t k = 1; k <= nz; ++k)
        for (int64_t j = 1; j <= ny; ++j)
            for (int64_t i = 1; i <= nx; ++i) {
                double kxx = iso4_deriv(S.cn_x, S, 0, i, j, k, false);
                double kyy = iso4_deriv(S.cn_y, S, 1, i, j, k, false);
                double kzz = iso4_deriv(S.cn_z, S, 2, i, j, k, false);
                double kxy = iso4_deriv(S.cn_x, S, 1, i, j, k, false);
                double kxz = iso4_deriv(S.cn_x, S, 2, i, j, k, false);
                double kyx = iso4_deriv(S.cn_y, S, 0, i, j, k, false);
                double kyz = iso4_deriv(S.cn_y, S, 2, i, j, k, false);
                double kzx = iso4_deriv(S.cn_z, S, 0, i, j, k, false);
                double kzy = iso4_deriv(S.cn_z, S, 1, i, j, k, false);
                int64_t n = S.s2(i, j, k);
                double cx = S.cn_x[n], cy = S.cn_y[n], cz = S.cn_z[n];
                S.curv[S.s1(i, j, k)] = (cx * cx - 1.0) * kxx + (cy * cy - 1.0) * kyy + (cz * cz - 1.0) * kzz +
                                        cx * cy * (kxy + kyx) + cx * cz * (kxz + kzx) + cy * cz * (kzy + kyz);
            }
}

// =====================================================================================
//  Velocity inlet (src/main_iteration_GPU.cu:1008-1077): bounce-back plus the momentum of
//  the inlet profile, per colour by the injected saturation.  Profile and mean velocity:
//  src/Init_multiphase.cpp:258-281, src/Misc.cpp:387-419 (1000-term duct series).
// =====================================================================================
void velocity_profile(OracleSim& S) {
    const int64_t nx = S.nx, ny = S.ny;
    S.w_in.assign(S.X1 * S.Y1, 0.0);
    const double U = S.p.uin_avg;
    for (int64_t y = 1; y <= ny; ++y)
        for (int64_t x = 1; x <= nx; ++x)
            if (x > 1 && x < nx && y > 1 && y < ny) S.w_in[S.col(x, y)] = U;
    const double la_x = double(nx - 1) - 0.5 - 0.5, la_y = double(ny - 1) - 0.5 - 0.5;
    const double a = 0.5 * la_x, b = 0.5 * la_y;
    double sum_t = 0.0;
    for (int64_t n = 1; n <= 1000; n += 2) sum_t += std::tanh(0.5 * double(n) * K.pi * b / a) / std::pow(double(n), 5);
    double c = 1.0 - 192.0 / std::pow(K.pi, 5) * (a / b) * sum_t;
    c = -3.0 * U / (c * std::pow(a, 2));
    for (int64_t y = 1; y <= ny; ++y)
        for (int64_t x = 1; x <= nx; ++x) {
            if (!(x > 1 && x < nx && y > 1 && y < ny)) continue;
            const double xx = x - 1.5 - a, yy = y - 1.5 - b;
            double acc = 0.0;
            for (int64_t n = 1; n <= 1000; n += 2) {
                acc += std::pow(-1.0, 0.5 * double(n - 1)) * std::cos(0.5 * double(n) * K.pi * xx / a) / std::pow(double(n), 3) *
                       (1.0 - (std::exp(0.5 * double(n) * K.pi * (yy - b) / a) + std::exp(0.5 * double(n) * K.pi * (-yy - b) / a)) /
                                  (1.0 + std::exp(0.5 * double(n) * K.pi * (-b - b) / a)));
            }
            S.w_in[S.col(x, y)] = acc * (-16.0 * c * std::pow(a, 2) * std::pow(K.pi, -3));
        }
}

void inlet_velocity(OracleSim& S, bool after_odd) {
    const double a1 = 6.0 * K.w1, a2 = 6.0 * K.w2;
    for (int64_t j = 1; j <= S.ny; ++j)
        for (int64_t i = 1; i <= S.nx; ++i) {
            const int w = S.walls[S.s2(i, j, 1)];
            double& p0 = S.phi[S.s4(i, j, 0)];
            p0 = S.p.phi_inlet * (1 - w) + p0 * w;
            for (int k = -1; k >= -3; --k) S.phi[S.s4(i, j, k)] = p0;
            double inj[2];
            inj[1] = S.w_in[S.col(i, j)] * S.p.relaxation;
            inj[0] = inj[1] * S.p.sa_inject;
            inj[1] = inj[1] - inj[0];
            for (int g = 0; g < 2; ++g) {
                auto f = [&](int64_t a, int64_t b, int64_t c, int e) -> double& { return S.F(a, b, c, e, g); };
                const double t = inj[g];
                if (!after_odd) {
                    f(i, j, 0, 5) = (f(i, j, 1, 6) + a1 * t) * (1 - w) + f(i, j, 0, 5) * w;
                    f(i - 1, j, 0, 11) = (f(i, j, 1, 14) + a2 * t) * (1 - w) + f(i - 1, j, 0, 11) * w;
                    f(i + 1, j, 0, 12) = (f(i, j, 1, 13) + a2 * t) * (1 - w) + f(i + 1, j, 0, 12) * w;
                    f(i, j - 1, 0, 15) = (f(i, j, 1, 18) + a2 * t) * (1 - w) + f(i, j - 1, 0, 15) * w;
                    f(i, j + 1, 0, 16) = (f(i, j, 1, 17) + a2 * t) * (1 - w) + f(i, j + 1, 0, 16) * w;
                } else {
                    f(i, j, 1, 6) = (f(i, j, 0, 5) + a1 * t) * (1 - w) + f(i, j, 1, 6) * w;
                    f(i, j, 1, 13) = (f(i + 1, j, 0, 12) + a2 * t) * (1 - w) + f(i, j, 1, 13) * w;
                    f(i, j, 1, 14) = (f(i - 1, j, 0, 11) + a2 * t) * (1 - w) + f(i, j, 1, 14) * w;
                    f(i, j, 1, 17) = (f(i, j + 1, 0, 16) + a2 * t) * (1 - w) + f(i, j, 1, 17) * w;
                    f(i, j, 1, 18) = (f(i, j - 1, 0, 15) + a2 * t) * (1 - w) + f(i, j, 1, 18) * w;
                }
            }
        }
}

// =====================================================================================
//  Convective outlet (src/main_iteration_GPU.cu:1245-1357): unknown populations and the
//  ghost phi advected out with the mean inlet velocity from the values of the last step.
// =====================================================================================
void outlet_convective(OracleSim& S, bool after_odd) {
    const int64_t nz = S.nz;
    const double u = S.p.uin_avg, r = 1.0 / (1.0 + u);
    for (int64_t j = 1; j <= S.ny; ++j)
        for (int64_t i = 1; i <= S.nx; ++i) {
            const int w = S.walls[S.s2(i, j, nz)];
            double& pg = S.phi[S.s4(i, j, nz + 1)];
            pg = ((S.CPHI(i, j) + u * S.phi[S.s4(i, j, nz)]) * r) * (1 - w) + pg * w;
            S.CPHI(i, j) = pg;
            for (int k = 2; k <= 4; ++k) S.phi[S.s4(i, j, nz + k)] = pg;
            // (stored slot, target cell, slot read one plane inside)
            struct Link { int e; int64_t a, b, c; int er; };
            const Link before[5] = {{6, i, j, nz + 1, 6}, {13, i - 1, j, nz + 1, 13}, {14, i + 1, j, nz + 1, 14},
                                    {17, i, j - 1, nz + 1, 17}, {18, i, j + 1, nz + 1, 18}};
            const Link after[5] = {{6, i, j, nz, 5}, {14, i, j, nz, 11}, {13, i, j, nz, 12},
                                   {18, i, j, nz, 15}, {17, i, j, nz, 16}};
            const Link* L = after_odd ? after : before;
            for (int g = 0; g < 2; ++g)
                for (int m = 0; m < 5; ++m) {
                    const Link& l = L[m];
                    double& t = S.F(l.a, l.b, l.c, l.er, g);
                    const double inner = S.F(l.a, l.b, l.c - 1, l.er, g);
                    t = ((S.CF(i, j, l.e, g) + u * inner) * r) * (1 - w) + t * w;
                }
            for (int g = 0; g < 2; ++g)
                for (int m = 0; m < 5; ++m) {
                    const Link& l = L[m];
                    S.CF(i, j, l.e, g) = S.F(l.a, l.b, l.c, l.er, g);
                }
        }
}

// state of a new run: the outlet plane of the initial populations and phi
// (src/Init_multiphase.cpp:445-494)
void convective_init(OracleSim& S) {
    S.conv.assign(39 * S.X1 * S.Y1, 0.0);
    for (int64_t j = 0; j <= S.ny + 1; ++j)
        for (int64_t i = 0; i <= S.nx + 1; ++i) {
            for (int g = 0; g < 2; ++g)
                for (int e = 0; e < 19; ++e) S.CF(i, j, e, g) = S.F(i, j, S.nz, e, g);
            S.CPHI(i, j) = S.phi[S.s4(i, j, S.nz)];
        }
}

// =====================================================================================
//  Porous plate at k = zp (src/main_iteration_GPU.cu:1744-1882): the blocked colour is
//  reflected on both faces of the plate, the other passes.
// =====================================================================================
void porous_plate(OracleSim& S, bool after_odd) {
    const int64_t zp = S.p.z_porous_plate;
    if (S.p.porous_plate == 0 || zp < 1 || zp > S.nz) return;
    const int b = S.p.porous_plate == 1 ? 0 : 1, c = 1 - b;
    const int up[5] = {6, 13, 14, 17, 18}, dn[5] = {5, 12, 11, 16, 15};
    // reflection pairs: (slot of the plate-plane cell, its lateral offset, slot read)
    const int rx[5] = {0, -1, 1, 0, 0}, ry[5] = {0, 0, 0, -1, 1};
    for (int64_t j = 1; j <= S.ny; ++j)
        for (int64_t i = 1; i <= S.nx; ++i) {
            if (!after_odd) {
                // f(i+dx, j+dy, zp, up) <- f(i, j, zp-1, dn); f(i-dx, j-dy, zp, dn) <- f(i, j, zp+1, up)
                for (int m = 0; m < 5; ++m) S.F(i + rx[m], j + ry[m], zp, up[m], b) = S.F(i, j, zp - 1, dn[m], b);
                for (int m = 0; m < 5; ++m) S.F(i - rx[m], j - ry[m], zp, dn[m], b) = S.F(i, j, zp + 1, up[m], b);
                for (int m = 0; m < 5; ++m) S.F(i, j, zp, up[m], c) = S.F(i, j, zp + 1, up[m], c);
                for (int m = 0; m < 5; ++m) S.F(i, j, zp, dn[m], c) = S.F(i, j, zp - 1, dn[m], c);
            } else {
                for (int m = 0; m < 5; ++m) S.F(i, j, zp - 1, dn[m], b) = S.F(i + rx[m], j + ry[m], zp, up[m], b);
                for (int m = 0; m < 5; ++m) S.F(i, j, zp + 1, up[m], b) = S.F(i - rx[m], j - ry[m], zp, dn[m], b);
                for (int m = 0; m < 5; ++m) S.F(i, j, zp - 1, dn[m], c) = S.F(i, j, zp, dn[m], c);
                for (int m = 0; m < 5; ++m) S.F(i, j, zp + 1, up[m], c) = S.F(i, j, zp, up[m], c);
            }
        }
}

void open_boundaries(OracleSim& S, bool odd) {
    if (S.open_z()) {
        if (S.p.inlet_bc == 1) inlet_velocity(S, odd);
        else if (S.p.inlet_bc == 2) inlet_zou_he(S, odd);
        if (S.p.outlet_bc == 1) outlet_convective(S, odd);
        else if (S.p.outlet_bc == 2) outlet_zou_he(S, odd);
    }
    porous_plate(S, odd);
}

// =====================================================================================
//  One time step (src/main_iteration_GPU.cu:1890-2055)
// =====================================================================================
void step(OracleSim& S, int ntime) {
    const bool odd = (ntime % 2) != 0;
    collide_sweep(S, odd);
    if (S.p.kper) { periodic_z(S, odd); periodic_phi(S, true, false); }
    if (S.p.jper) { periodic_y(S, odd); periodic_phi(S, false, true); }
    if (S.p.jper && S.p.kper) { periodic_zy(S, odd); periodic_phi(S, true, true, true); }
    open_boundaries(S, odd);
    gradient_chain(S);
}

}  // namespace

// =====================================================================================
//  C ABI
// =====================================================================================
extern "C" {

void* oracle_create(const oracle_params* p, const int8_t* geometry) {
    if (!p || !geometry) return nullptr;
    OracleSim* S = new OracleSim();
    S->p = *p;
    S->nx = p->nx; S->ny = p->ny; S->nz = p->nz;
    S->X1 = S->nx + 2; S->Y1 = S->ny + 2; S->Z1 = S->nz + 2;
    S->X2 = S->nx + 4; S->Y2 = S->ny + 4; S->Z2 = S->nz + 4;
    S->X4 = S->nx + 8; S->Y4 = S->ny + 8; S->Z4 = S->nz + 8;
    S->S1 = S->X1 * S->Y1 * S->Z1; S->S2 = S->X2 * S->Y2 * S->Z2; S->S4 = S->X4 * S->Y4 * S->Z4;
    S->walls.assign(S->S2, 0);
    S->walls_type.assign(S->S4, 0);
    S->s_nx.assign(S->S4, 0.0); S->s_ny.assign(S->S4, 0.0); S->s_nz.assign(S->S4, 0.0);
    S->pdf.assign(38 * S->S1, 0.0);
    S->phi.assign(S->S4, 0.0);
    S->cn_x.assign(S->S2, 0.0); S->cn_y.assign(S->S2, 0.0); S->cn_z.assign(S->S2, 0.0); S->c_norm.assign(S->S2, 0.0);
    S->curv.assign(S->S1, 0.0);
    preprocess_geometry(*S, geometry);
    if (S->open_z() && p->inlet_bc == 1) velocity_profile(*S);
    if (S->open_z() && p->outlet_bc == 1) S->conv.assign(39 * S->X1 * S->Y1, 0.0);
    return S;
}

void oracle_destroy(void* h) { delete static_cast<OracleSim*>(h); }

int oracle_init_phi_random(void* h, double sa_target, uint64_t seed) {
    init_phi(*static_cast<OracleSim*>(h), 6, 0.0, sa_target, seed);
    return 0;
}

double oracle_node_uniform(uint64_t seed, uint64_t n) { return node_uniform(seed, n); }

int oracle_init_phi(void* h, int option, double interface_z0) {
    if (option < 1 || option > 5) return -1;  // option 6: oracle_init_phi_random
    init_phi(*static_cast<OracleSim*>(h), option, interface_z0);
    return 0;
}

void oracle_init_pdf(void* h, const double* rho, const double* u, const double* v, const double* w) {
    OracleSim& S = *static_cast<OracleSim*>(h);
    init_pdf(S, rho, u, v, w);
    if (S.open_z() && S.p.outlet_bc == 1) convective_init(S);
}

void oracle_init_pdf_relaxed(void* h, const double* rho, const double* u, const double* v, const double* w) {
    OracleSim& S = *static_cast<OracleSim*>(h);
    init_pdf(S, rho, u, v, w, true);
    if (S.open_z() && S.p.outlet_bc == 1) convective_init(S);
}

void oracle_color_gradient(void* h) { gradient_chain(*static_cast<OracleSim*>(h)); }

void oracle_step(void* h, int ntime) { step(*static_cast<OracleSim*>(h), ntime); }

void oracle_run(void* h, int ntime0, int nsteps) {
    OracleSim& S = *static_cast<OracleSim*>(h);
    for (int t = 0; t < nsteps; ++t) step(S, ntime0 + t);
}

// Sub-step entry points used by the pinning tests.
void oracle_collide(void* h, int odd) { collide_sweep(*static_cast<OracleSim*>(h), odd != 0); }
void oracle_boundaries(void* h, int odd) {
    OracleSim& S = *static_cast<OracleSim*>(h);
    if (S.p.kper) { periodic_z(S, odd); periodic_phi(S, true, false); }
    if (S.p.jper) { periodic_y(S, odd); periodic_phi(S, false, true); }
    if (S.p.jper && S.p.kper) { periodic_zy(S, odd); periodic_phi(S, true, true, true); }
    open_boundaries(S, odd);
}

// Collision of a single node with explicit inputs (for the MRT pinning test).
double oracle_collide_node(const oracle_params* p, const double* g1, const double* g2, double cnx, double cny,
                           double cnz, double curv, double cnorm, double* out1, double* out2) {
    return collide_node(*p, g1, g2, cnx, cny, cnz, curv, cnorm, out1, out2);
}

// Array access: 0 pdf, 1 phi, 2 cn_x, 3 cn_y, 4 cn_z, 5 c_norm, 6 curv, 7 s_nx, 8 s_ny, 9 s_nz,
// 10 w_in (X1 * Y1, velocity inlet), 11 conv (39 * X1 * Y1, convective outlet)
void* oracle_array(void* h, int which, int64_t* count) {
    OracleSim& S = *static_cast<OracleSim*>(h);
    std::vector<double>* v = nullptr;
    switch (which) {
        case 0: v = &S.pdf; break;
        case 1: v = &S.phi; break;
        case 2: v = &S.cn_x; break;
        case 3: v = &S.cn_y; break;
        case 4: v = &S.cn_z; break;
        case 5: v = &S.c_norm; break;
        case 6: v = &S.curv; break;
        case 7: v = &S.s_nx; break;
        case 8: v = &S.s_ny; break;
        case 9: v = &S.s_nz; break;
        case 10: v = &S.w_in; break;
        case 11: v = &S.conv; break;
        case 12: v = &S.smooth; break;
        default: return nullptr;
    }
    if (count) *count = (int64_t)v->size();
    return v->data();
}
// 0 walls (s2), 1 walls_type (s4), 2 pore_profile_z (nz)
int32_t* oracle_int_array(void* h, int which, int64_t* count) {
    OracleSim& S = *static_cast<OracleSim*>(h);
    std::vector<int32_t>* v = which == 0 ? &S.walls : which == 1 ? &S.walls_type : which == 2 ? &S.pore_profile_z : nullptr;
    if (!v) return nullptr;
    if (count) *count = (int64_t)v->size();
    return v->data();
}

// compute_macro_vars (src/Misc.cpp:222-274): rho, u, v, w on s1 interior; phi zeroed at
// solids like the reference's host copy.
void oracle_macro(void* h, double* rho, double* u, double* v, double* w, int zero_solid_phi) {
    OracleSim& S = *static_cast<OracleSim*>(h);
    const int64_t nx = S.nx, ny = S.ny, nz = S.nz;
#pragma omp parallel for collapse(2) schedule(static)
    for (int64_t k = 1; k <= nz; ++k)
        for (int64_t j = 1; j <= ny; ++j)
            for (int64_t i = 1; i <= nx; ++i) {
                const int wi = S.walls[S.s2(i, j, k)];
                const int64_t n = S.s1(i, j, k), n2 = S.s2(i, j, k);
                double ft[19];
                for (int q = 0; q < 19; ++q) ft[q] = S.pdf[n + S.S1 * q] + S.pdf[n + S.S1 * (q + 19)];
                double r = ft[0];
                for (int q = 1; q < 19; ++q) r = r + ft[q];
                rho[n] = r * (1 - wi);
                double tmp = 0.5 * S.p.lbm_gamma * S.curv[n] * S.c_norm[n2];
                double fx = tmp * S.cn_x[n2], fy = tmp * S.cn_y[n2], fz = tmp * S.cn_z[n2] + S.p.force_z;
                u[n] = (ft[1] - ft[2] + ft[7] - ft[8] + ft[9] - ft[10] + ft[11] - ft[12] + ft[13] - ft[14] - 0.5 * fx) * (1 - wi);
                v[n] = (ft[3] - ft[4] + ft[7] + ft[8] - ft[9] - ft[10] + ft[15] - ft[16] + ft[17] - ft[18] - 0.5 * fy) * (1 - wi);
                w[n] = (ft[5] - ft[6] + ft[11] + ft[12] - ft[13] - ft[14] + ft[15] + ft[16] - ft[17] - ft[18] - 0.5 * fz) * (1 - wi);
                if (zero_solid_phi) {
                    double& ph = S.phi[S.s4(i, j, k)];
                    ph = 0.0 * wi + ph * (1 - wi);
                }
            }
}

// Monitor sums of src/Monitor.cpp:17-80, accumulated serially in the reference's order.
// rho/u/v/w must come from oracle_macro of the same state; phi as stored.
void oracle_monitor_sums(void* h, const double* rho, const double* u, const double* v, const double* w,
                         oracle_monitor* out) {
    OracleSim& S = *static_cast<OracleSim*>(h);
    const int64_t nx = S.nx, ny = S.ny, nz = S.nz;
    double umax = 0.0, usq1 = 0.0, usq2 = 0.0;
    for (int64_t k = 1; k <= nz; ++k)
        for (int64_t j = 1; j <= ny; ++j)
            for (int64_t i = 1; i <= nx; ++i) {
                const int wi = S.walls[S.s2(i, j, k)];
                const int64_t n = S.s1(i, j, k);
                double t = (u[n] * u[n] + v[n] * v[n] + w[n] * w[n]) * (1 - wi);
                if (umax < t) umax = t;
                const double ph = S.phi[S.s4(i, j, k)];
                if (ph > 0.999) usq1 = usq1 + t;
                else if (ph < -0.999) usq2 = usq2 + t;
            }
    for (int64_t k = 1; k <= nz; ++k) {
        double t1 = 0, t2 = 0, pk = 0, t3 = 0, t4 = 0, t5 = 0, t6 = 0;
        for (int64_t j = 1; j <= ny; ++j)
            for (int64_t i = 1; i <= nx; ++i) {
                const int wi = S.walls[S.s2(i, j, k)];
                const int64_t n = S.s1(i, j, k);
                const double ph = S.phi[S.s4(i, j, k)];
                t3 = t3 + 0.5 * (1.0 + ph) * (1 - wi);
                t4 = t4 + 0.5 * (1.0 - ph) * (1 - wi);
                t5 = t5 + rho[n] * 0.5 * (1.0 + ph) * (1 - wi);
                t6 = t6 + rho[n] * 0.5 * (1.0 - ph) * (1 - wi);
                t1 = t1 + w[n] * 0.5 * (1.0 + ph) * (1 - wi);
                t2 = t2 + w[n] * 0.5 * (1.0 - ph) * (1 - wi);
                pk = pk + rho[n] * (1 - wi);
            }
        out->pre[k - 1] = pk; out->fl1[k - 1] = t1; out->fl2[k - 1] = t2;
        out->vol1[k - 1] = t3; out->vol2[k - 1] = t4; out->mass1[k - 1] = t5; out->mass2[k - 1] = t6;
    }
    out->umax2 = umax; out->usq1 = usq1; out->usq2 = usq2;
}

int64_t oracle_count(void* h, int which) {
    OracleSim& S = *static_cast<OracleSim*>(h);
    return which == 0 ? S.n_solid_boundary : S.n_fluid_boundary;
}

int oracle_threads() { return omp_get_max_threads(); }

}  // extern "C"
