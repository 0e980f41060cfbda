// Geometry preprocessing on the device: the reference's host routine
// geometry_preprocessing_new (src/Geometry_preprocessing.cpp:29-403) as five kernels.
//   k_geo_fill      ghost layers of the wall flags (10 layers; periodic or copy rule)
//   k_geo_classify  walls_type: -1 fluid boundary, 0 fluid, 1 solid, 2 solid boundary
//   k_geo_smooth    one Jacobi sweep of the D3Q27 filter (run 4 times)
//   k_geo_normals   ISO8 gradient of the smoothed flags at fluid-boundary nodes
//   k_geo_extract   walls (s2, 0/1) and walls_type (s4) for the step kernels
// Built with -ffp-contract=off: the normals feed the wetting model and must round like
// the FP64 oracle.
#include <hip/hip_runtime.h>
#include <cmath>
#include <cstdint>
#include <cfloat>
#include <utility>
#include "geometry.h"
#include "lattice.h"

namespace mflbm {
namespace geo {

constexpr int G = 10;  // ghost layers of the temporary arrays (src/Geometry_preprocessing.cpp:37)

struct TDims {
    int nx, ny, nz, TX, TY, TZ;
    int64_t N;
};

__device__ __forceinline__ int64_t ti(const TDims& t, int i, int j, int k) {
    return (int64_t)(i + G - 1) + t.TX * ((int64_t)(j + G - 1) + (int64_t)t.TY * (k + G - 1));
}

// Source coordinate of a ghost cell after the reference's sequential fills (z, then y,
// then x; src/Geometry_preprocessing.cpp:56-123).  A periodic lower ghost whose image
// lies below the first layer reads a cell the loop has not written yet, i.e. 0.
__device__ __forceinline__ int gmap(int i, int n, int per, bool& zero) {
    if (i >= 1 && i <= n) return i;
    if (!per) return i < 1 ? 1 : n;
    if (i < 1) {
        const int s = n + i;
        if (s < 1) { zero = true; return 1; }
        return s;
    }
    int s = i - n;
    while (s > n) s -= n;
    return s;
}

__global__ void k_geo_fill(const int8_t* __restrict__ geo, int8_t* __restrict__ wt, TDims t, int iper, int jper,
                           int kper) {
    const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= t.N) return;
    const int i = (int)(n % t.TX) - G + 1, j = (int)((n / t.TX) % t.TY) - G + 1, k = (int)(n / ((int64_t)t.TX * t.TY)) - G + 1;
    bool zero = false;
    const int si = gmap(i, t.nx, iper, zero), sj = gmap(j, t.ny, jper, zero), sk = gmap(k, t.nz, kper, zero);
    wt[n] = zero ? 0 : geo[(si - 1) + (int64_t)t.nx * ((sj - 1) + (int64_t)t.ny * (sk - 1))];
}

__global__ void k_geo_classify(const int8_t* __restrict__ wt, int8_t* __restrict__ wc, TDims t) {
    const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= t.N) return;
    const int i = (int)(n % t.TX) - G + 1, j = (int)((n / t.TX) % t.TY) - G + 1, k = (int)(n / ((int64_t)t.TX * t.TY)) - G + 1;
    int8_t v = wt[n];
    const bool inner = i >= 2 - G && i <= t.nx + G - 1 && j >= 2 - G && j <= t.ny + G - 1 && k >= 2 - G && k <= t.nz + G - 1;
    if (inner) {
        if (v == 1) {
            for (int q = 1; q <= 18; ++q)
                if (wt[ti(t, i + EX[q], j + EY[q], k + EZ[q])] <= 0) { v = 2; break; }
        } else if (v == 0) {
            for (int q = 1; q <= 18; ++q)
                if (wt[ti(t, i + EX[q], j + EY[q], k + EZ[q])] >= 1) { v = -1; break; }
        }
    }
    wc[n] = v;
}

__global__ void k_geo_tofloat(const int8_t* __restrict__ wt, double* __restrict__ a, double* __restrict__ b, int64_t N) {
    const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    a[n] = (double)wt[n];
    b[n] = (double)wt[n];
}

// 27-point order of src/Geometry_preprocessing.cpp:31-33
__constant__ int SX[27] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1};
__constant__ int SY[27] = {0, 0, 0, 1, -1, 0, 0, 1, 1, -1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, -1, 1, -1, 1};
__constant__ int SZ[27] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, 1, -1, -1, 1, 1, -1, -1, -1, 1, 1, -1, -1, 1, 1, -1};

__global__ void k_geo_smooth(const double* __restrict__ in, double* __restrict__ out, TDims t) {
    const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= t.N) return;
    const int i = (int)(n % t.TX) - G + 1, j = (int)((n / t.TX) % t.TY) - G + 1, k = (int)(n / ((int64_t)t.TX * t.TY)) - G + 1;
    if (!(i >= 2 - G && i <= t.nx + G - 1 && j >= 2 - G && j <= t.ny + G - 1 && k >= 2 - G && k <= t.nz + G - 1)) return;
    const double we[4] = {8.0 / 27.0, 2.0 / 27.0, 1.0 / 54.0, 1.0 / 216.0};
    double acc = 0.0;
    for (int m = 0; m < 27; ++m) {
        const int w2 = SX[m] * SX[m] + SY[m] * SY[m] + SZ[m] * SZ[m];
        acc += in[ti(t, i + SX[m], j + SY[m], k + SZ[m])] * we[w2];
    }
    out[n] = acc;
}

// ISO8 stencil plus-offsets per axis in the reference's term order
// (src/Geometry_preprocessing.cpp:232-384); groups of 1,4,4,1,8,12,4 terms.
__constant__ int ISO8_OFF[3][34][3] = {
    {{1, 0, 0},
     {1, 1, 0}, {1, -1, 0}, {1, 0, 1}, {1, 0, -1},
     {1, 1, 1}, {1, 1, -1}, {1, -1, 1}, {1, -1, -1},
     {2, 0, 0},
     {2, 1, 0}, {2, -1, 0}, {2, 0, 1}, {2, 0, -1}, {1, 2, 0}, {1, -2, 0}, {1, 0, 2}, {1, 0, -2},
     {2, 1, 1}, {2, 1, -1}, {2, -1, 1}, {2, -1, -1}, {1, 2, 1}, {1, 2, -1}, {1, -2, 1}, {1, -2, -1},
     {1, 1, 2}, {1, 1, -2}, {1, -1, 2}, {1, -1, -2},
     {2, 2, 0}, {2, -2, 0}, {2, 0, 2}, {2, 0, -2}},
    {{0, 1, 0},
     {1, 1, 0}, {-1, 1, 0}, {0, 1, 1}, {0, 1, -1},
     {1, 1, 1}, {1, 1, -1}, {-1, 1, -1}, {-1, 1, 1},
     {0, 2, 0},
     {2, 1, 0}, {-2, 1, 0}, {0, 2, 1}, {0, 2, -1}, {1, 2, 0}, {-1, 2, 0}, {0, 1, 2}, {0, 1, -2},
     {2, 1, 1}, {2, 1, -1}, {-2, 1, 1}, {-2, 1, -1}, {1, 2, 1}, {1, 2, -1}, {-1, 2, 1}, {-1, 2, -1},
     {1, 1, 2}, {1, 1, -2}, {-1, 1, 2}, {-1, 1, -2},
     {2, 2, 0}, {-2, 2, 0}, {0, 2, 2}, {0, 2, -2}},
    {{0, 0, 1},
     {0, 1, 1}, {0, -1, 1}, {1, 0, 1}, {-1, 0, 1},
     {1, 1, 1}, {1, -1, 1}, {-1, 1, 1}, {-1, -1, 1},
     {0, 0, 2},
     {0, 1, 2}, {0, -1, 2}, {2, 0, 1}, {-2, 0, 1}, {0, 2, 1}, {0, -2, 1}, {1, 0, 2}, {-1, 0, 2},
     {2, 1, 1}, {2, -1, 1}, {-2, 1, 1}, {-2, -1, 1}, {1, 2, 1}, {1, -2, 1}, {-1, 2, 1}, {-1, -2, 1},
     {1, 1, 2}, {1, -1, 2}, {-1, 1, 2}, {-1, -1, 2},
     {0, 2, 2}, {0, -2, 2}, {2, 0, 2}, {-2, 0, 2}}};
__constant__ int ISO8_CNT[7] = {1, 4, 4, 1, 8, 12, 4};

__global__ void k_geo_normals(const int8_t* __restrict__ wc, const double* __restrict__ sm, TDims t, Dims d,
                              double* __restrict__ snx, double* __restrict__ sny, double* __restrict__ snz) {
    const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= d.S4) return;
    const int i = (int)(n % d.X4) - 3, j = (int)((n / d.X4) % d.Y4) - 3, k = (int)(n / ((int64_t)d.X4 * d.Y4)) - 3;
    double out[3] = {0.0, 0.0, 0.0};
    if (wc[ti(t, i, j, k)] == -1) {
        const double w8[7] = {4.0 / 45.0, 1.0 / 21.0, 2.0 / 105.0, 5.0 / 504.0, 1.0 / 315.0, 1.0 / 630.0, 1.0 / 5040.0};
        double nw[3];
        for (int a = 0; a < 3; ++a) {
            double tot = 0.0;
            int e = 0;
            for (int g = 0; g < 7; ++g) {
                double acc = 0.0;
                for (int m = 0; m < ISO8_CNT[g]; ++m, ++e) {
                    const int ox = ISO8_OFF[a][e][0], oy = ISO8_OFF[a][e][1], oz = ISO8_OFF[a][e][2];
                    const double pl = sm[ti(t, i + ox, j + oy, k + oz)];
                    const double mi = sm[ti(t, i - ox, j - oy, k - oz)];
                    acc = (m == 0) ? pl : acc + pl;
                    acc = acc - mi;
                }
                const double term = w8[g] * acc;
                tot = (g == 0) ? term : tot + term;
            }
            nw[a] = tot;
        }
        const double tmp = 1.0 / (sqrt(nw[0] * nw[0] + nw[1] * nw[1] + nw[2] * nw[2]) + (double)FLT_EPSILON);
        out[0] = nw[0] * tmp;
        out[1] = nw[1] * tmp;
        out[2] = nw[2] * tmp;
    }
    snx[n] = out[0];
    sny[n] = out[1];
    snz[n] = out[2];
}

__global__ void k_geo_extract(const int8_t* __restrict__ wt, const int8_t* __restrict__ wc, TDims t, Dims d,
                              uint8_t* __restrict__ walls, int8_t* __restrict__ wtype) {
    const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (n < d.S4) {
        const int i = (int)(n % d.X4) - 3, j = (int)((n / d.X4) % d.Y4) - 3, k = (int)(n / ((int64_t)d.X4 * d.Y4)) - 3;
        wtype[n] = wc[ti(t, i, j, k)];
    }
    if (n < d.S2) {
        const int i = (int)(n % d.X2) - 1, j = (int)((n / d.X2) % d.Y2) - 1, k = (int)(n / ((int64_t)d.X2 * d.Y2)) - 1;
        walls[n] = (uint8_t)wt[ti(t, i, j, k)];
    }
}

}  // namespace geo

static inline unsigned nblk(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

hipError_t geometry_preprocess(const Dims& d, int iper, int jper, int kper, const int8_t* geo_dev, uint8_t* walls,
                               int8_t* wtype, double* snx, double* sny, double* snz, hipStream_t s) {
    using namespace geo;
    TDims t;
    t.nx = d.nx; t.ny = d.ny; t.nz = d.nz;
    t.TX = d.nx + 2 * G; t.TY = d.ny + 2 * G; t.TZ = d.nz + 2 * G;
    t.N = (int64_t)t.TX * t.TY * t.TZ;
    int8_t *wt = nullptr, *wc = nullptr;
    double *a = nullptr, *b = nullptr;
    hipError_t e;
    if ((e = hipMalloc(&wt, t.N)) != hipSuccess) return e;
    if ((e = hipMalloc(&wc, t.N)) != hipSuccess) { (void)hipFree(wt); return e; }
    if ((e = hipMalloc(&a, t.N * sizeof(double))) != hipSuccess) { (void)hipFree(wt); (void)hipFree(wc); return e; }
    if ((e = hipMalloc(&b, t.N * sizeof(double))) != hipSuccess) { (void)hipFree(wt); (void)hipFree(wc); (void)hipFree(a); return e; }
    const int B = 256;
    k_geo_fill<<<nblk(t.N, B), B, 0, s>>>(geo_dev, wt, t, iper, jper, kper);
    k_geo_classify<<<nblk(t.N, B), B, 0, s>>>(wt, wc, t);
    k_geo_tofloat<<<nblk(t.N, B), B, 0, s>>>(wt, a, b, t.N);
    for (int it = 0; it < 4; ++it) {
        k_geo_smooth<<<nblk(t.N, B), B, 0, s>>>(a, b, t);
        std::swap(a, b);
    }
    // after the swaps `a` holds the last sweep (the reference's walls_smooth_2)
    k_geo_normals<<<nblk(d.S4, B), B, 0, s>>>(wc, a, t, d, snx, sny, snz);
    k_geo_extract<<<nblk(d.S4 > d.S2 ? d.S4 : d.S2, B), B, 0, s>>>(wt, wc, t, d, walls, wtype);
    e = hipGetLastError();
    hipError_t e2 = hipStreamSynchronize(s);
    (void)hipFree(wt); (void)hipFree(wc); (void)hipFree(a); (void)hipFree(b);
    return e != hipSuccess ? e : e2;
}

}  // namespace mflbm
