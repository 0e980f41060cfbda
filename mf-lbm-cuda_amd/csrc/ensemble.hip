// Batched single-phase ensemble (mflbm_ensemble_* in include/mflbm.h).
//
// n independent domains share one set of slot arrays: the fluid nodes of every domain, each
// domain padded to whole 256-node blocks, then the bounce-back cells of all domains.  Periodic
// y and z are resolved in the neighbour table itself (a wrapped neighbour is the fluid node
// it wraps to), so a step is one collide launch over all domains and nothing else.  A block
// belongs to one domain; a domain whose velocity field is steady is switched off by a flag
// its blocks read first.
//
// The collision is the colour-gradient collision with phi = 1 (mrt_collide, the path the
// FP32 two-colour kernel takes where fluid 2 is absent) on FP32 deviations from the rest
// state rho_d (4/21, 13/126, 1/63), rho_d the domain's relaxed-start density.  Starting at
// rest, every deviation is zero.
#include <hip/hip_runtime.h>
#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "../../include/mflbm.h"
#include "device_lock.h"
#include "kernels.inl"
#include "layout.h"

using namespace mflbm;

namespace {

constexpr int EB = 256;  // threads per block; each domain's fluid range is padded to it

struct EnsDev {
    float* f;                  // 19 slot arrays of ns cells: deviations from the rest state
    int64_t ns;                // cells per slot array
    int64_t nf;                // padded fluid cells of all domains (multiple of EB)
    const int32_t* nb;         // 18 * nf: cell of (x + e_q) in slot array opp(q)
    const uint32_t* nb16;      // the same links as 16-bit rows against per-wave bases (Sparse<T>)
    const int32_t* nbb;
    const int32_t* blk_dom;    // domain of each block
    const int32_t* dom_start;  // first fluid cell of each domain
    const int32_t* dom_count;  // fluid nodes of each domain
    const int32_t* active;     // 1 while the domain is being stepped
    float nui, force;
    int mrt;                   // collision model (mflbm_params::mrt)
};

// AA step of one fluid node (the access pattern of k_collide, one colour).
template <bool ODD>
__global__ __launch_bounds__(EB) void k_ens_collide(EnsDev E) {
    const int b = dev::xcd_block();
    const int d = E.blk_dom[b];
    if (!E.active[d]) return;
    const int64_t t = (int64_t)b * EB + threadIdx.x;
    if (t >= (int64_t)E.dom_start[d] + E.dom_count[d]) return;
    const int64_t NS = E.ns, NF = E.nf;
    float ft[19];
    int64_t nb[19];
    if (ODD) {
        nb[0] = t;
        dev::odd_links(E.nbb, E.nb16, E.nb, NF, t, nb);
    }
#pragma unroll
    for (int q = 0; q < 19; ++q) {
        const int64_t src = ODD ? nb[OPP[q]] : t;
        const int slot = ODD ? q : OPP[q];
        // even: own cells, streamed once -> nontemporal (as k_collide)
        ft[q] = ODD ? E.f[slot * NS + src] : __builtin_nontemporal_load(E.f + slot * NS + src);
    }
    float m_rho = ft[0];
#pragma unroll
    for (int q = 1; q < 19; ++q) m_rho = m_rho + ft[q];
    // harmonic viscosity of collide_node at phi = 1
    const float omega = 1.0f / (6.0f / ((1.0f + 1.0f) * E.nui + (1.0f - 1.0f) * E.nui) + 0.5f);
    float f[19];
    dev::mrt_collide<float>(ft, m_rho, 0.0f, 0.0f, E.force, omega, f, E.mrt);
#pragma unroll
    for (int q = 0; q < 19; ++q) {
        const int64_t dst = ODD ? nb[q] : t;
        const int slot = ODD ? OPP[q] : q;
        if (ODD) E.f[slot * NS + dst] = f[q];
        else __builtin_nontemporal_store(f[q], E.f + slot * NS + dst);
    }
}

// After an even step: macro velocity (compute_macro_vars: j - F/2) and density deviation,
// the velocity's change since the last check, and per-block FP64 sums of |du|^2, |u|^2
// and w.
__global__ __launch_bounds__(EB) void k_ens_check(EnsDev E, float* __restrict__ ux, float* __restrict__ uy,
                                                  float* __restrict__ uz, float* __restrict__ drho,
                                                  double* __restrict__ part) {
    const int b = blockIdx.x;
    const int d = E.blk_dom[b];
    const int64_t t = (int64_t)b * EB + threadIdx.x;
    double acc[3] = {0.0, 0.0, 0.0};
    if (E.active[d] && t < (int64_t)E.dom_start[d] + E.dom_count[d]) {
        float ft[19];
#pragma unroll
        for (int q = 0; q < 19; ++q) ft[q] = E.f[q * E.ns + t];
        const float u = ft[1] - ft[2] + ft[7] - ft[8] + ft[9] - ft[10] + ft[11] - ft[12] + ft[13] - ft[14] - 0.5f * 0.0f;
        const float v = ft[3] - ft[4] + ft[7] + ft[8] - ft[9] - ft[10] + ft[15] - ft[16] + ft[17] - ft[18] - 0.5f * 0.0f;
        const float w = ft[5] - ft[6] + ft[11] + ft[12] - ft[13] - ft[14] + ft[15] + ft[16] - ft[17] - ft[18] -
                        0.5f * E.force;
        const double du = (double)u - ux[t], dv = (double)v - uy[t], dw = (double)w - uz[t];
        acc[0] = du * du + dv * dv + dw * dw;
        acc[1] = (double)u * u + (double)v * v + (double)w * w;
        acc[2] = w;
        ux[t] = u;
        uy[t] = v;
        uz[t] = w;
        float r = ft[0];
#pragma unroll
        for (int q = 1; q < 19; ++q) r = r + ft[q];
        drho[t] = r;
    }
    double s;
    if (dev::block_tree<3, EB>(acc, s)) part[(int64_t)b * 3 + threadIdx.x] = s;
}

// One block per domain: its blocks' partial sums in a fixed order.
__global__ __launch_bounds__(EB) void k_ens_reduce(const double* __restrict__ part, const int32_t* __restrict__ dom_start,
                                                   const int32_t* __restrict__ dom_count, double* __restrict__ out) {
    const int d = blockIdx.x;
    const int64_t b0 = dom_start[d] / EB, nb = (dom_count[d] + EB - 1) / EB;
    double acc[3] = {0.0, 0.0, 0.0};
    for (int64_t b = threadIdx.x; b < nb; b += EB)
#pragma unroll
        for (int m = 0; m < 3; ++m) acc[m] += part[(b0 + b) * 3 + m];
    double s;
    if (dev::block_tree<3, EB>(acc, s)) out[(int64_t)d * 3 + threadIdx.x] = s;
}

// Scatter the stored fields of domains first .. first+count-1 into dense arrays.
__global__ __launch_bounds__(EB) void k_ens_scatter(EnsDev E, const float* __restrict__ ux, const float* __restrict__ uy,
                                                    const float* __restrict__ uz, const float* __restrict__ drho,
                                                    const int32_t* __restrict__ loc, int64_t cell0, int64_t ncell,
                                                    int first, int64_t npt, float* __restrict__ out) {
    const int64_t t = cell0 + (int64_t)blockIdx.x * EB + threadIdx.x;
    if (t >= cell0 + ncell) return;
    const int d = E.blk_dom[t / EB];
    if (t >= (int64_t)E.dom_start[d] + E.dom_count[d]) return;
    float* o = out + (int64_t)(d - first) * 4 * npt + loc[t];
    o[0] = ux[t];
    o[npt] = uy[t];
    o[2 * npt] = uz[t];
    o[3 * npt] = drho[t];
}

// Initial populations of the fluid nodes of domains first .. from dense fields (u, v, w and
// rho - rho_d per node, the layout of mflbm_ensemble_fields): the collision's rest weights
// times rho plus the reference's velocity terms, stored as deviations from rho_d r_q (the
// FP64 arithmetic of k_init_eq's relaxed FP32 branch).
__global__ __launch_bounds__(EB) void k_ens_init(EnsDev E, const double* __restrict__ rho_ref,
                                                 const int32_t* __restrict__ loc, int64_t cell0, int64_t ncell,
                                                 int first, int64_t npt, const float* __restrict__ in) {
    const int64_t t = cell0 + (int64_t)blockIdx.x * EB + threadIdx.x;
    if (t >= cell0 + ncell) return;
    const int d = E.blk_dom[t / EB];
    if (t >= (int64_t)E.dom_start[d] + E.dom_count[d]) return;
    const float* x = in + (int64_t)(d - first) * 4 * npt + loc[t];
    const double uu = x[0], vv = x[npt], ww = x[2 * npt], rd = rho_ref[d], rr = rd + (double)x[3 * npt];
    const double usq = uu * uu + vv * vv + ww * ww;
    const double cu[19] = {0.0, uu, -uu, vv, -vv, ww, -ww, uu + vv, -uu + vv, uu - vv, -uu - vv,
                           uu + ww, -uu + ww, uu - ww, -uu - ww, vv + ww, -vv + ww, vv - ww, -vv - ww};
#pragma unroll
    for (int q = 0; q < 19; ++q) {
        const double wqq = q == 0 ? 1.0 / 3.0 : (q <= 6 ? 1.0 / 18.0 : 1.0 / 36.0);
        const double shape = wqq * (3.0 * cu[q] + 4.5 * cu[q] * cu[q] - 1.5 * usq);
        E.f[q * E.ns + t] = (float)((rr - rd) * rest_weight(q) + rr * shape);
    }
}

}  // namespace

struct mflbm_ensemble {
    int nd = 0;
    int64_t nx = 0, ny = 0, nz = 0, npt = 0;
    double nu = 0.0, force = 0.0;
    int mrt = 2;
    int64_t nf = 0, ns = 0, nf_real = 0;
    std::vector<int32_t> start, count;
    float* f = nullptr;
    uint32_t* nb16 = nullptr;
    int32_t* nbb = nullptr;
    int n_escaped = 0;  // waves on the 32-bit rows
    int32_t *nb = nullptr, *blk_dom = nullptr, *dom_start = nullptr, *dom_count = nullptr, *active = nullptr;
    int32_t* loc = nullptr;
    float *ux = nullptr, *uy = nullptr, *uz = nullptr, *drho = nullptr;
    double *part = nullptr, *red = nullptr;
    hipStream_t stream = nullptr;
    hipGraphExec_t graph = nullptr;
    int graph_steps = 0;
    int32_t ntime = 1;  // next step number (odd at every check boundary)
    // per-domain progress, kept across solve calls
    std::vector<int32_t> taken, checks;
    std::vector<double> perm, change;
    std::vector<double> rho_d;  // rest density of each domain
    std::string err;
};

namespace {

int fail(mflbm_ensemble* e, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    if (e) e->err = buf;
    return -1;
}

#define ECHK(e, x)                                                                          \
    do {                                                                                    \
        hipError_t _r = (x);                                                                \
        if (_r != hipSuccess) return fail(e, "%s: %s", #x, hipGetErrorString(_r));          \
    } while (0)

EnsDev dev_view(const mflbm_ensemble* e) {
    EnsDev E;
    E.f = e->f; E.ns = e->ns; E.nf = e->nf; E.nb = e->nb; E.nb16 = e->nb16; E.nbb = e->nbb; E.blk_dom = e->blk_dom;
    E.dom_start = e->dom_start; E.dom_count = e->dom_count; E.active = e->active;
    E.nui = (float)(1.0 / e->nu);
    E.force = (float)e->force;
    E.mrt = e->mrt;
    return E;
}

void free_all(mflbm_ensemble* e) {
    if (e->graph) (void)hipGraphExecDestroy(e->graph);
    e->graph = nullptr;
    void* p[] = {e->f, e->nb, e->nb16, e->nbb, e->blk_dom, e->dom_start, e->dom_count, e->active, e->loc, e->ux, e->uy, e->uz,
                 e->drho, e->part, e->red};
    for (void* x : p)
        if (x) (void)hipFree(x);
    e->f = nullptr; e->nb16 = nullptr; e->nbb = nullptr;
    e->nb = e->blk_dom = e->dom_start = e->dom_count = e->active = e->loc = nullptr;
    e->ux = e->uy = e->uz = e->drho = nullptr; e->part = e->red = nullptr;
    if (e->stream) (void)hipStreamDestroy(e->stream);
    e->stream = nullptr;
}

// Per-domain host work: solid mask (x faces walls), fluid count, bounce-back cells per slot
// and links per direction.
struct DomainCounts {
    int32_t nf = 0;
    int64_t ext[19] = {0};
    int64_t links[19] = {0};  // fluid x whose neighbour x - e_q is solid
};

void count_domain(const int8_t* geo, int64_t nx, int64_t ny, int64_t nz, DomainCounts& c) {
    auto solid = [&](int64_t i, int64_t j, int64_t k) {
        return i == 0 || i == nx - 1 || geo[i + nx * (j + ny * k)] != 0;
    };
    for (int64_t k = 0; k < nz; ++k)
        for (int64_t j = 0; j < ny; ++j)
            for (int64_t i = 0; i < nx; ++i) {
                if (solid(i, j, k)) continue;
                c.nf++;
                for (int q = 1; q < 19; ++q) {
                    const int64_t jj = (j + EY[q] + ny) % ny, kk = (k + EZ[q] + nz) % nz;
                    if (solid(i + EX[q], jj, kk)) {
                        c.ext[OPP[q]]++;
                        c.links[OPP[q]]++;
                    }
                }
            }
}

}  // namespace

extern "C" {

int mflbm_ensemble_create(const mflbm_params* params, const int8_t* geometries, int32_t n_domains, double* rho_ref,
                          mflbm_ensemble** out) {
    std::lock_guard<std::recursive_mutex> lock(device_lock());
    if (!out) return -1;
    *out = nullptr;
    auto* e = new mflbm_ensemble();
    *out = e;
    if (!params || !geometries || n_domains <= 0) return fail(e, "ensemble: null argument or no domains");
    if (params->nx < 3 || params->ny < 1 || params->nz < 1) return fail(e, "ensemble: domain too small");
    if (!(params->la_nui1 > 0.0)) return fail(e, "ensemble: la_nui1 must be positive");
    e->nd = n_domains;
    e->nx = params->nx; e->ny = params->ny; e->nz = params->nz;
    e->npt = e->nx * e->ny * e->nz;
    e->nu = 1.0 / params->la_nui1;
    e->force = params->force_z;
    e->mrt = params->mrt ? params->mrt : 2;
    const int64_t npt = e->npt, nx = e->nx, ny = e->ny, nz = e->nz;
    const int nd = n_domains;

    unsigned nth = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    auto parallel = [&](auto&& fn) {
        std::vector<std::thread> th;
        for (unsigned w = 0; w < nth; ++w)
            th.emplace_back([&, w] {
                for (int d = (int)w; d < nd; d += (int)nth) fn(d);
            });
        for (auto& x : th) x.join();
    };

    // pass 1: counts, offsets and rest densities
    std::vector<DomainCounts> cnt(nd);
    parallel([&](int d) { count_domain(geometries + (int64_t)d * npt, nx, ny, nz, cnt[d]); });
    e->start.resize(nd);
    e->count.resize(nd);
    std::vector<int64_t> ext_base((size_t)nd * 19);
    int64_t nf = 0;
    for (int d = 0; d < nd; ++d) {
        e->start[d] = (int32_t)nf;
        e->count[d] = cnt[d].nf;
        e->nf_real += cnt[d].nf;
        nf += (cnt[d].nf + EB - 1) / EB * EB;
        if (nf >= INT32_MAX) return fail(e, "ensemble: too many fluid nodes for 32-bit cell indices");
    }
    if (nf == 0) return fail(e, "ensemble: no fluid node in any domain");
    int64_t ext_max = 0;
    for (int s = 0; s < 19; ++s) {
        int64_t run = nf;
        for (int d = 0; d < nd; ++d) {
            ext_base[(size_t)d * 19 + s] = run;
            run += cnt[d].ext[s];
        }
        ext_max = std::max(ext_max, run - nf);
    }
    e->nf = nf;
    e->ns = (nf + ext_max + 1023) / 1024 * 1024;
    if (e->ns >= INT32_MAX) return fail(e, "ensemble: too many cells for 32-bit cell indices");
    e->rho_d.assign(nd, 1.0);
    {
        for (int d = 0; d < nd; ++d) {
            // the cold start's mass (D3Q19 weights in every bounce-back link) at the relaxed
            // weights (the summation order of relaxed_mean_density in mflbm/warmstart.py)
            double w_std = 0.0, r_rel = 0.0;
            for (int q = 1; q < 19; ++q) {
                const double links = (double)cnt[d].links[q];
                w_std += links * (q <= 6 ? 1.0 / 18.0 : 1.0 / 36.0);
                r_rel += links * (q <= 6 ? 13.0 / 126.0 : 1.0 / 63.0);
            }
            e->rho_d[d] = ((double)cnt[d].nf + w_std) / ((double)cnt[d].nf + r_rel);
            if (rho_ref) rho_ref[d] = e->rho_d[d];
        }
    }

    // pass 2: neighbour rows, node coordinates and block owners
    std::vector<int32_t> nb((size_t)18 * nf, 0), loc((size_t)nf, 0), blk((size_t)(nf / EB), 0);
    parallel([&](int d) {
        const int8_t* geo = geometries + (int64_t)d * npt;
        auto solid = [&](int64_t i, int64_t j, int64_t k) {
            return i == 0 || i == nx - 1 || geo[i + nx * (j + ny * k)] != 0;
        };
        std::vector<int32_t> idx((size_t)npt, -1);
        int32_t t = 0;
        for (int64_t n = 0; n < npt; ++n) {
            const int64_t i = n % nx, j = (n / nx) % ny, k = n / (nx * ny);
            if (!solid(i, j, k)) idx[n] = t++;
        }
        const int64_t base = e->start[d];
        int64_t ext_next[19];
        for (int s = 0; s < 19; ++s) ext_next[s] = ext_base[(size_t)d * 19 + s];
        for (int64_t n = 0; n < npt; ++n) {
            const int32_t tl = idx[n];
            if (tl < 0) continue;
            const int64_t i = n % nx, j = (n / nx) % ny, k = n / (nx * ny);
            loc[base + tl] = (int32_t)n;
            for (int q = 1; q < 19; ++q) {
                const int64_t ii = i + EX[q], jj = (j + EY[q] + ny) % ny, kk = (k + EZ[q] + nz) % nz;
                const int64_t m = ii + nx * (jj + ny * kk);
                const int64_t cell = idx[m] >= 0 ? base + idx[m] : ext_next[OPP[q]]++;
                nb[(size_t)(q - 1) * nf + base + tl] = (int32_t)cell;
            }
        }
        const int64_t nblk = (cnt[d].nf + EB - 1) / EB;
        for (int64_t b = 0; b < nblk; ++b) blk[(size_t)(base / EB + b)] = d;
    });
    // short neighbour rows, domain by domain (a wave never spans two domains: their ranges
    // are padded to whole blocks); the padding nodes are not encoded
    std::vector<uint32_t> nb16((size_t)9 * nf, 0u);
    std::vector<int32_t> nbb((size_t)(nf / 64) * NB_WAVE, 0);
    std::vector<int> esc(nd, 0);
    parallel([&](int d) {
        const int64_t b = e->start[d], end = b + e->count[d];
        const int64_t w0 = b / 64, w1 = (b + (e->count[d] + EB - 1) / EB * EB) / 64;
        esc[d] = encode_short_rows(nb.data(), nf, w0, w1, [&](int64_t t) { return t < end; }, nb16.data(), nbb.data());
    });
    for (int d = 0; d < nd; ++d) {
        if (esc[d] < 0) return fail(e, "ensemble: short neighbour row of domain %d does not decode", d);
        e->n_escaped += esc[d];
    }

    ECHK(e, hipStreamCreate(&e->stream));
    auto up = [&](int32_t** dst, const std::vector<int32_t>& v) -> hipError_t {
        hipError_t r = hipMalloc(dst, v.size() * sizeof(int32_t));
        if (r != hipSuccess) return r;
        return hipMemcpy(*dst, v.data(), v.size() * sizeof(int32_t), hipMemcpyHostToDevice);
    };
    ECHK(e, up(&e->nb, nb));
    ECHK(e, up(&e->nbb, nbb));
    ECHK(e, hipMalloc(&e->nb16, nb16.size() * sizeof(uint32_t)));
    ECHK(e, hipMemcpy(e->nb16, nb16.data(), nb16.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
    ECHK(e, up(&e->loc, loc));
    ECHK(e, up(&e->blk_dom, blk));
    ECHK(e, up(&e->dom_start, e->start));
    ECHK(e, up(&e->dom_count, e->count));
    std::vector<int32_t> act(nd);
    for (int d = 0; d < nd; ++d) act[d] = e->count[d] > 0;
    ECHK(e, up(&e->active, act));
    ECHK(e, hipMalloc(&e->f, (size_t)19 * e->ns * sizeof(float)));
    ECHK(e, hipMemset(e->f, 0, (size_t)19 * e->ns * sizeof(float)));  // rest state: zero deviations
    for (float** u : {&e->ux, &e->uy, &e->uz, &e->drho}) {
        ECHK(e, hipMalloc(u, (size_t)nf * sizeof(float)));
        ECHK(e, hipMemset(*u, 0, (size_t)nf * sizeof(float)));
    }
    ECHK(e, hipMalloc(&e->part, (size_t)(nf / EB) * 3 * sizeof(double)));
    ECHK(e, hipMalloc(&e->red, (size_t)nd * 3 * sizeof(double)));
    e->taken.assign(nd, 0);
    e->checks.assign(nd, 0);
    e->perm.assign(nd, 0.0);
    e->change.assign(nd, 1.0);
    return 0;
}

void mflbm_ensemble_destroy(mflbm_ensemble* e) {
    std::lock_guard<std::recursive_mutex> lock(device_lock());
    if (!e) return;
    free_all(e);
    delete e;
}

const char* mflbm_ensemble_last_error(const mflbm_ensemble* e) { return e ? e->err.c_str() : "null ensemble"; }

int64_t mflbm_ensemble_fluid_nodes(const mflbm_ensemble* e) { return e ? e->nf_real : -1; }

int mflbm_ensemble_solve(mflbm_ensemble* e, int32_t check, double tol, int32_t max_steps, int32_t* steps,
                         double* permeability, double* change, double* device_ms) {
    if (!e || !e->f) return -1;
    if (check <= 0 || check % 2 != 0) return fail(e, "ensemble: check must be a positive even number of steps");
    if (max_steps < 0) return fail(e, "ensemble: negative max_steps");
    const int nd = e->nd;
    const EnsDev E = dev_view(e);
    const int nblk = (int)(e->nf / EB);
    if (!e->graph || e->graph_steps != check) {
        if (e->graph) (void)hipGraphExecDestroy(e->graph);
        e->graph = nullptr;
        hipGraph_t g = nullptr;
        ECHK(e, hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal));
        for (int s = 0; s < check; ++s) {
            if (s % 2 == 0) k_ens_collide<true><<<nblk, EB, 0, e->stream>>>(E);
            else k_ens_collide<false><<<nblk, EB, 0, e->stream>>>(E);
        }
        ECHK(e, hipStreamEndCapture(e->stream, &g));
        hipError_t r = hipGraphInstantiate(&e->graph, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        ECHK(e, r);
        e->graph_steps = check;
    }
    std::vector<int32_t> act(nd);
    ECHK(e, hipMemcpy(act.data(), e->active, nd * sizeof(int32_t), hipMemcpyDeviceToHost));
    std::vector<double> red((size_t)nd * 3);
    hipEvent_t ev0, ev1;
    ECHK(e, hipEventCreate(&ev0));
    ECHK(e, hipEventCreate(&ev1));
    double ms_total = 0.0;
    const double scale = e->nu / ((double)e->npt * e->force);
    int taken = 0;
    bool any = std::any_of(act.begin(), act.end(), [](int a) { return a != 0; });
    while (any && taken + check <= max_steps) {
        ECHK(e, hipEventRecord(ev0, e->stream));
        ECHK(e, hipGraphLaunch(e->graph, e->stream));
        ECHK(e, hipEventRecord(ev1, e->stream));
        k_ens_check<<<nblk, EB, 0, e->stream>>>(E, e->ux, e->uy, e->uz, e->drho, e->part);
        k_ens_reduce<<<nd, EB, 0, e->stream>>>(e->part, e->dom_start, e->dom_count, e->red);
        ECHK(e, hipGetLastError());
        ECHK(e, hipMemcpyAsync(red.data(), e->red, red.size() * sizeof(double), hipMemcpyDeviceToHost, e->stream));
        ECHK(e, hipStreamSynchronize(e->stream));
        float ms = 0.0f;
        ECHK(e, hipEventElapsedTime(&ms, ev0, ev1));
        ms_total += ms;
        e->ntime += check;
        taken += check;
        bool changed = false;
        any = false;
        for (int d = 0; d < nd; ++d) {
            if (!act[d]) continue;
            e->taken[d] += check;
            e->perm[d] = red[(size_t)d * 3 + 2] * scale;
            const double uu = red[(size_t)d * 3 + 1];
            e->change[d] = uu > 0.0 ? std::sqrt(red[(size_t)d * 3]) / std::sqrt(uu) : 1.0;
            // the first check has no earlier field to compare with
            if (e->checks[d]++ > 0 && e->change[d] < tol) {
                act[d] = 0;
                changed = true;
            } else {
                any = true;
            }
        }
        if (changed) ECHK(e, hipMemcpy(e->active, act.data(), nd * sizeof(int32_t), hipMemcpyHostToDevice));
    }
    (void)hipEventDestroy(ev0);
    (void)hipEventDestroy(ev1);
    for (int d = 0; d < nd; ++d) {
        if (steps) steps[d] = act[d] ? -e->taken[d] : e->taken[d];
        if (permeability) permeability[d] = e->perm[d];
        if (change) change[d] = e->change[d];
    }
    if (device_ms) *device_ms = ms_total;
    return 0;
}

int mflbm_ensemble_init_fields(mflbm_ensemble* e, int32_t first, int32_t count, const float* fields) {
    if (!e || !e->f || !fields) return -1;
    if (first < 0 || count <= 0 || first + count > e->nd) return fail(e, "ensemble: domain range out of bounds");
    for (int d = first; d < first + count; ++d)
        if (e->taken[d] > 0) return fail(e, "ensemble: initial fields only before a domain's first step");
    const int64_t cell0 = e->start[first];
    const int64_t cell1 = (int64_t)e->start[first + count - 1] + (e->count[first + count - 1] + EB - 1) / EB * EB;
    const int64_t ncell = cell1 - cell0;
    const size_t bytes = (size_t)count * 4 * e->npt * sizeof(float);
    float* buf = nullptr;
    double* rref = nullptr;
    ECHK(e, hipMalloc(&buf, bytes));
    hipError_t r = hipMalloc(&rref, (size_t)e->nd * sizeof(double));
    if (r == hipSuccess) r = hipMemcpy(buf, fields, bytes, hipMemcpyHostToDevice);
    if (r == hipSuccess) r = hipMemcpy(rref, e->rho_d.data(), (size_t)e->nd * sizeof(double), hipMemcpyHostToDevice);
    const EnsDev E = dev_view(e);
    if (r == hipSuccess && ncell > 0) {
        k_ens_init<<<(unsigned)((ncell + EB - 1) / EB), EB, 0, e->stream>>>(E, rref, e->loc, cell0, ncell, first,
                                                                           e->npt, buf);
        r = hipGetLastError();
    }
    // the steady test of the first check compares with this state
    if (r == hipSuccess) {
        k_ens_check<<<(unsigned)(e->nf / EB), EB, 0, e->stream>>>(E, e->ux, e->uy, e->uz, e->drho, e->part);
        r = hipGetLastError();
    }
    if (r == hipSuccess) r = hipStreamSynchronize(e->stream);
    (void)hipFree(buf);
    (void)hipFree(rref);
    if (r != hipSuccess) return fail(e, "ensemble init: %s", hipGetErrorString(r));
    for (int d = first; d < first + count; ++d) e->checks[d] = 1;
    return 0;
}

int mflbm_ensemble_fields(mflbm_ensemble* e, int32_t first, int32_t count, float* out) {
    if (!e || !e->f || !out) return -1;
    if (first < 0 || count <= 0 || first + count > e->nd) return fail(e, "ensemble: domain range out of bounds");
    const int64_t cell0 = e->start[first];
    const int64_t cell1 = (int64_t)e->start[first + count - 1] + (e->count[first + count - 1] + EB - 1) / EB * EB;
    const size_t bytes = (size_t)count * 4 * e->npt * sizeof(float);
    float* buf = nullptr;
    ECHK(e, hipMalloc(&buf, bytes));
    hipError_t r = hipMemsetAsync(buf, 0, bytes, e->stream);
    const int64_t ncell = cell1 - cell0;
    if (r == hipSuccess && ncell > 0) {
        k_ens_scatter<<<(unsigned)((ncell + EB - 1) / EB), EB, 0, e->stream>>>(dev_view(e), e->ux, e->uy, e->uz,
                                                                                e->drho, e->loc, cell0, ncell, first,
                                                                                e->npt, buf);
        r = hipGetLastError();
    }
    if (r == hipSuccess) r = hipStreamSynchronize(e->stream);
    if (r == hipSuccess) r = hipMemcpy(out, buf, bytes, hipMemcpyDeviceToHost);
    (void)hipFree(buf);
    if (r != hipSuccess) return fail(e, "ensemble fields: %s", hipGetErrorString(r));
    return 0;
}

}  // extern "C"
