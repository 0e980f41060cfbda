// Fused colour-gradient chain: phi extrapolation to solid-boundary nodes (K3), interface
// normals with the wetting correction (K4+K5), normal extrapolation to solid-boundary
// nodes (K6) and curvature (K7), producing the force records of the next collision in one
// pass over phi (src/main_iteration_GPU.cu:732-1003 run as four dependent kernels there).
//
// One workgroup owns a TX x TY column of interior nodes over a run of `kchunk` z-planes and
// marches up z.  Each stage needs its input one node further out than its output, so the
// loaded planes carry a 4-node halo in x and y, and the stages trail the load front:
//   load    plane p     phi, tag                          (TX+8) x (TY+8)
//   K3      plane p-1   phi at solid-boundary nodes       (TX+6) x (TY+6)
//   K4+K5   plane p-2   normals + c_norm at fluid nodes   (TX+4) x (TY+4)
//   K6      plane p-3   normals at solid-boundary nodes   (TX+2) x (TY+2)
//   K7      plane p-4   curvature -> force record         TX x TY
// Ring buffers in LDS keep the planes still in use.  K3 and K6 work in place: they write
// only solid-boundary entries and read only fluid entries, and the D3Q19 neighbours of a
// fluid node are fluid or solid-boundary nodes by construction of walls_type.
//
// Node types never reach LDS: each thread keeps the tag words of the entries it loaded
// (the same entries it classifies into the stage lists) for the last four planes, and
// what a stage needs of them travels in its list entry -- the fluid-neighbour mask of a
// solid-boundary node (K3, K6) and the wetting flag of a fluid node (K4).
//
// Every value goes through the same device functions as the split kernels (iso4_at,
// interface_normal, curvature_of, the K3/K6 loops in q order), so both paths round alike.
// The phi written back is the solid-boundary phi of the owned planes (plus the ghost plane
// k = 0 for the first z-run and k = nz + 1 for the last): the only solid-boundary values a
// later kernel of the step reads (the open inlets and outlets, which keep the ghost phi of
// wall columns, and the periodic copies).  The full dense gradient
// fields are produced on request by the split kernels (they are functions of phi alone).
#pragma once
#include <type_traits>
#include "kernels.inl"

namespace mflbm {
namespace dev {


// v if f, else +0.0: a bit mask, so that the compiler cannot turn the weighted sums into
// branches around their loads (it does for a select).
__device__ __forceinline__ double keep_if(bool f, double v) {
    return __longlong_as_double(__double_as_longlong(v) & (f ? ~0ll : 0ll));
}
__device__ __forceinline__ float keep_if(bool f, float v) { return __int_as_float(__float_as_int(v) & (f ? ~0 : 0)); }

// Four values of one node side by side, so that one vector access moves all four (16 B in
// FP32, 32 B in FP64) and a row of nodes is one contiguous run: the two-pass chain's
// (cn_x, cn_y, cn_z, c_norm) and its solid normals (snx, sny, snz, 0) of wetting nodes.
template <typename T>
struct alignas(4 * sizeof(T)) Vec4 {
    T x, y, z, w;
};

// Workgroup barrier for data exchanged through LDS only.  __syncthreads() also waits for
// every outstanding global load of the wave (its release fence covers all memory), which
// would expose the latency of the next plane's loads at every stage boundary.
__device__ __forceinline__ void lds_barrier() { asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory"); }

template <typename T, int TX, int TY, int GRAD_THREADS>
__global__ __launch_bounds__(GRAD_THREADS) void k_gradient(GradArgs<T> A, StepConst<T> P) {
    constexpr int GRAD_WAVES = GRAD_THREADS / 64;
    constexpr int H = 4;
    constexpr int WX = TX + 2 * H, WY = TY + 2 * H, NP = WX * WY;  // loaded plane
    constexpr int VX = TX + 4, VY = TY + 4, NV = VX * VY;          // normals plane (halo 2)
    constexpr int NO = TX * TY;                                    // owned plane
    constexpr int SEG = (NP + GRAD_WAVES - 1) / GRAD_WAVES;        // entries per wave
    constexpr int NSUB = (SEG + 63) / 64;                          // entries per lane
    static_assert(NP <= 1024, "list entries keep the plane position in 10 bits");
    constexpr int N3 = (TX + 6) * (TY + 6), N4L = NV, N6 = (TX + 2) * (TY + 2);  // list capacities
    __shared__ T s_phi[4][NP];
    __shared__ T s_cn[4][3][NV];
    __shared__ T s_cnorm[3][NO];
    // work lists of the four stages, double-buffered by iteration parity: the loaded-plane
    // entries a stage computes, packed in one run per list (each wave reserves its place
    // with one LDS atomic on the list's count).  K3 and K6 entries carry the fluid-neighbour
    // mask (e | mask << 10), K4 entries the wetting flag (e | wet << 14), K7 entries come
    // with the node's force-record index.
    __shared__ int32_t s_l3[2][N3], s_l6[2][N6];
    __shared__ int16_t s_l4[2][N4L], s_l7[2][NO];
    __shared__ int32_t s_r7[2][NO];
    __shared__ unsigned long long s_cnt[2];  // list counts (16 bits each: K3, K4, K6, K7)
    // 1 when K4 found a non-zero normal anywhere in the normals region of the plane
    __shared__ int s_any[4];
    // K3/K6 weight sums: the D3Q19 weights of the fluid neighbours summed in q order are the
    // 1/18 of the a axis neighbours added a times, then the 1/36 of the d diagonal ones d
    // times, so the sum depends on (a, d) only; s_ws[a][d] holds it, summed that way
    __shared__ T s_ws[7][13];

    const Dims& d = A.d;
    const int b = xcd_block();
    const int tx = b % A.tiles_x, ty = (b / A.tiles_x) % A.tiles_y, tz = b / (A.tiles_x * A.tiles_y);
    const int ib = 1 + tx * TX - H, jb = 1 + ty * TY - H;  // node of loaded-plane entry (0, 0)
    const int k0 = 1 + tz * A.kchunk, k1 = min(d.nz, k0 + A.kchunk - 1);
    const int kw0 = tz == 0 ? 0 : k0, kw1 = k1 == d.nz ? d.nz + 1 : k1;
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int nxo = min(TX, d.nx - (ib + H) + 1), nyo = min(TY, d.ny - (jb + H) + 1);  // owned extent

    auto s4 = [](int p) { return (p + 8) & 3; };
    auto s3 = [](int p) { return (p + 12) % 3; };

    // Entry of lane `lane`, sub-step sb of this wave: the mapping of the loads and of the
    // list classification alike, so that a thread classifies the entries it loaded.
    auto entry = [&](int sb, bool& in) {
        const int j = sb * 64 + lane, e = wave * SEG + j;
        in = j < SEG && e < NP;
        return e;
    };
    // register staging of the planes two ahead of the store: two register sets, the loop
    // unrolled by two, so that a plane's loads have two iterations of stage work to arrive
    // in (kept raw until the store, so that nothing waits for them early).  Against one
    // plane ahead: 0.680 -> 0.652 ms at 256^3 FP64, 0.546 -> 0.496 ms in FP32.
    T pvb[2][NSUB];
    int32_t tvb[2][NSUB];
    auto fetch = [&](int p, T (&pv)[NSUB], int32_t (&tv)[NSUB]) {
#pragma unroll
        for (int sb = 0; sb < NSUB; ++sb) {
            bool in;
            const int e = entry(sb, in);
            const int gi = ib + e % WX, gj = jb + e / WX;
            if (in && gi <= d.nx + 4 && gj <= d.ny + 4) {
                const int c = i4(d, gi, gj, p);
                pv[sb] = A.phi[c];
                tv[sb] = A.tag[c];
            } else {
                pv[sb] = T(0);
                tv[sb] = TAG_SOLID;
            }
        }
    };
    // tag words of this thread's entries in planes p-1 .. p-4 (th[0] = p-1)
    int32_t th[4][NSUB];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int sb = 0; sb < NSUB; ++sb) th[a][sb] = TAG_SOLID;
    // plane position of a list entry, clamped so that no address formed from it can leave
    // the arrays
    auto pos = [](int v) { return min(v & 1023, NP - 1); };
    auto count = [&](int pr, int l) { return (int)((s_cnt[pr] >> (16 * l)) & 0xFFFF); };
    auto wsum = [&](uint32_t mask) -> T { return s_ws[__popc(mask & 63u)][__popc(mask >> 6)]; };

    if (tid < 7 * 13) {
        const int a = tid / 13, dd = tid % 13;
        T ws = T(0);
        for (int m = 0; m < a; ++m) ws += W<T>::w1;
        for (int m = 0; m < dd; ++m) ws += W<T>::w2;
        s_ws[a][dd] = ws;
    }
    if (tid < 2) s_cnt[tid] = 0;
    __syncthreads();

    fetch(k0 - H, pvb[0], tvb[0]);
    if (k0 - H + 1 <= k1 + H) fetch(k0 - H + 1, pvb[1], tvb[1]);
    int32_t tcur[NSUB];
#pragma unroll
    for (int sb = 0; sb < NSUB; ++sb) tcur[sb] = TAG_SOLID;
    auto body = [&](int p, auto B) {
        constexpr int b = decltype(B)::value;
        T (&pv)[NSUB] = pvb[b];
        int32_t (&tv)[NSUB] = tvb[b];
        const int par = p & 1;
#pragma unroll
        for (int sb = 0; sb < NSUB; ++sb) {
            th[3][sb] = th[2][sb];
            th[2][sb] = th[1][sb];
            th[1][sb] = th[0][sb];
            th[0][sb] = tcur[sb];
            tcur[sb] = tv[sb];
        }
        {
            T* ph = s_phi[s4(p)];
#pragma unroll
            for (int sb = 0; sb < NSUB; ++sb) {
                bool in;
                const int e = entry(sb, in);
                if (in) ph[e] = pv[sb];
            }
        }
        if (p + 2 <= k1 + H) fetch(p + 2, pv, tv);
        // the normals slot K4 fills in this iteration held plane p-6, which no stage reads
        // any more; its flag starts over
        if (tid == 0) s_any[s4(p - 2)] = 0;
        // stage lists of this iteration, from the tags of earlier planes kept in registers:
        // each wave counts its entries, reserves a run of each list with one LDS atomic on
        // the four packed counts and writes its entries there (the order within a list does
        // not matter: every entry is computed on its own)
        {
            auto flags = [&](int sb, bool& f3, bool& f4, bool& f6, bool& f7) {
                bool in;
                const int e = entry(sb, in);
                const int x = e % WX, y = e / WX;
                const bool r3 = in && x >= 1 && x < WX - 1 && y >= 1 && y < WY - 1;
                const bool r2 = in && x >= 2 && x < WX - 2 && y >= 2 && y < WY - 2;
                const bool r1 = in && x >= 3 && x < WX - 3 && y >= 3 && y < WY - 3;
                const bool r0 = in && x >= H && x < H + nxo && y >= H && y < H + nyo;
                f3 = r3 && (th[0][sb] & 3) == TAG_SOLID_BND;
                f4 = r2 && tag_fluid(th[1][sb] & 3);
                f6 = r1 && (th[2][sb] & 3) == TAG_SOLID_BND;
                f7 = r0 && tag_fluid(th[3][sb] & 3);
                return e;
            };
            uint64_t cnt = 0;  // four 16-bit counts: K3, K4, K6, K7
#pragma unroll
            for (int sb = 0; sb < NSUB; ++sb) {
                bool f3, f4, f6, f7;
                flags(sb, f3, f4, f6, f7);
                cnt += (uint64_t)__popcll(__ballot(f3)) | (uint64_t)__popcll(__ballot(f4)) << 16 |
                       (uint64_t)__popcll(__ballot(f6)) << 32 | (uint64_t)__popcll(__ballot(f7)) << 48;
            }
            uint64_t base = 0;
            if (lane == 0) base = atomicAdd(&s_cnt[par], (unsigned long long)cnt);
            base = ((uint64_t)__builtin_amdgcn_readfirstlane((int)(base >> 32)) << 32) |
                   (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)base);
            int b3 = (int)(base & 0xFFFF), b4 = (int)((base >> 16) & 0xFFFF);
            int b6 = (int)((base >> 32) & 0xFFFF), b7 = (int)(base >> 48);
            auto below = [&](uint64_t m) {
                return (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0));
            };
#pragma unroll
            for (int sb = 0; sb < NSUB; ++sb) {
                bool f3, f4, f6, f7;
                const int e = flags(sb, f3, f4, f6, f7);
                const uint64_t m3 = __ballot(f3), m4 = __ballot(f4), m6 = __ballot(f6), m7 = __ballot(f7);
                if (f3) s_l3[par][b3 + below(m3)] = e | ((th[0][sb] >> 2) << 10);
                if (f4) s_l4[par][b4 + below(m4)] = (int16_t)(e | (((th[1][sb] & 3) == TAG_WET) << 14));
                if (f6) s_l6[par][b6 + below(m6)] = e | ((th[2][sb] >> 2) << 10);
                if (f7) {
                    const int at = b7 + below(m7);
                    s_l7[par][at] = (int16_t)e;
                    s_r7[par][at] = th[3][sb] >> 2;
                }
                b3 += __popcll(m3); b4 += __popcll(m4); b6 += __popcll(m6); b7 += __popcll(m7);
            }
        }
        lds_barrier();
        // the lists of the other parity were last read before this barrier (by K7 of the
        // previous iteration); their counts start over for the next iteration
        if (tid == 0) s_cnt[par ^ 1] = 0;

        // Global reads of the K4 and K7 stages, issued now so that their latency hides
        // behind K3: the solid normals of the fluid-boundary nodes of plane p-2 and the
        // force-record index of the owned fluid nodes of plane p-4.
        constexpr int N4 = (NV + GRAD_THREADS - 1) / GRAD_THREADS;
        T wn[N4][3];
        int32_t rec = 0;
        if (p - 2 >= k0 - 2) {
            const int q = p - 2;
            const int16_t* list = s_l4[par];
            const int n = min(count(par, 1), N4L);
#pragma unroll
            for (int m = 0; m < N4; ++m) {
                const int i = tid + m * GRAD_THREADS;
                wn[m][0] = wn[m][1] = wn[m][2] = T(0);
                if (i < n) {
                    const int v = list[i], e = pos(v);
                    if (v & (1 << 14)) {
                        // three scalar loads: the 4-vector copy measured no faster here (FP64)
                        const int c = i4(d, ib + e % WX, jb + e / WX, q);
                        wn[m][0] = A.snx[c]; wn[m][1] = A.sny[c]; wn[m][2] = A.snz[c];
                    }
                }
            }
        }
        if (p - 4 >= k0 && tid < min(count(par, 3), NO)) rec = s_r7[par][tid];

        // K3: phi at solid-boundary nodes of plane q (src/main_iteration_GPU.cu:732-755)
        if (p - 1 >= k0 - 3) {
            const int q = p - 1;
            T* pc = s_phi[s4(q)];
            const T* pl[3] = {s_phi[s4(q - 1)], pc, s_phi[s4(q + 1)]};
            const bool wb = q >= kw0 && q <= kw1;
            const int32_t* list = s_l3[par];
            const int n = min(count(par, 0), N3);
            for (int i = tid; i < n; i += GRAD_THREADS) {
                const int lv = list[i], e = pos(lv), x = e % WX, y = e / WX;
                const uint32_t mask = (uint32_t)lv >> 10;
                // branch-free sum: a non-fluid neighbour adds +0 * w = +0, which leaves the
                // (never negative-zero) partial sum unchanged
                T ps = T(0);
#pragma unroll
                for (int qq = 1; qq < 19; ++qq) {
                    const int o = e + EX[qq] + WX * EY[qq];
                    const bool f = (mask >> (qq - 1)) & 1u;
                    ps += keep_if(f, pl[1 + EZ[qq]][o]) * wq<T>(qq);
                }
                const T v = ps / wsum(mask);
                pc[e] = v;
                if (wb && x >= H && x < H + nxo && y >= H && y < H + nyo) A.phi[i4(d, ib + x, jb + y, q)] = v;
            }
        }
        lds_barrier();

        // K4 + K5: normals and c_norm at fluid nodes of plane q (src/main_iteration_GPU.cu:757-878)
        if (p - 2 >= k0 - 2) {
            const int q = p - 2;
            const T* pl[3] = {s_phi[s4(q - 1)], s_phi[s4(q)], s_phi[s4(q + 1)]};
            T* cn = &s_cn[s4(q)][0][0];
            T* cno = s_cnorm[s3(q)];
            const int16_t* list = s_l4[par];
            const int n = min(count(par, 1), N4L);
#pragma unroll
            for (int m = 0; m < N4; ++m) {
                const int i = tid + m * GRAD_THREADS;
                if (i >= n) break;
                const int v = list[i], e = pos(v), x = e % WX, y = e / WX;
                const int r = (x - 2) + VX * (y - 2);
                auto a = [&](int dx, int dy, int dz) -> T { return pl[1 + dz][e + dx + WX * dy]; };
                const T cx = iso4_at<T, 0>(a);
                const T cy = iso4_at<T, 1>(a);
                const T cz = iso4_at<T, 2>(a);
                T vx, vy, vz;
                const T cnorm = interface_normal<T>(
                    cx, cy, cz, (v & (1 << 14)) != 0, P.cos_theta,
                    [&](T& nwx, T& nwy, T& nwz) { nwx = wn[m][0]; nwy = wn[m][1]; nwz = wn[m][2]; },
                    vx, vy, vz);
                cn[r] = vx; cn[NV + r] = vy; cn[2 * NV + r] = vz;
                if (x >= H && x < H + TX && y >= H && y < H + TY) cno[(x - H) + TX * (y - H)] = cnorm;
                if (cnorm != T(0)) s_any[s4(q)] = 1;
            }
        }
        lds_barrier();

        // K6: normals at solid-boundary nodes of plane q (src/main_iteration_GPU.cu:880-906)
        if (p - 3 >= k0 - 1) {
            const int q = p - 3;
            const T* cl[3] = {&s_cn[s4(q - 1)][0][0], &s_cn[s4(q)][0][0], &s_cn[s4(q + 1)][0][0]};
            T* cn = &s_cn[s4(q)][0][0];
            const int32_t* list = s_l6[par];
            const int n = min(count(par, 2), N6);
            // with no interface normal in planes q-1 .. q+1 every average below is +0 / ws
            const bool any = (s_any[s4(q - 1)] | s_any[s4(q)] | s_any[s4(q + 1)]) != 0;
            for (int i = tid; i < n; i += GRAD_THREADS) {
                const int lv = list[i], e = pos(lv), x = e % WX, y = e / WX;
                const uint32_t mask = (uint32_t)lv >> 10;
                const int v = (x - 2) + VX * (y - 2);
                if (!any) {
                    cn[v] = T(0); cn[NV + v] = T(0); cn[2 * NV + v] = T(0);
                    continue;
                }
                T sx = T(0), sy = T(0), sz = T(0);
#pragma unroll
                for (int qq = 1; qq < 19; ++qq) {
                    const bool f = (mask >> (qq - 1)) & 1u;
                    const T* c = cl[1 + EZ[qq]];
                    const int o = v + EX[qq] + VX * EY[qq];
                    sx += keep_if(f, c[o]) * wq<T>(qq);
                    sy += keep_if(f, c[NV + o]) * wq<T>(qq);
                    sz += keep_if(f, c[2 * NV + o]) * wq<T>(qq);
                }
                const T ws = wsum(mask);
                cn[v] = sx / ws; cn[NV + v] = sy / ws; cn[2 * NV + v] = sz / ws;
            }
        }
        lds_barrier();

        // K7: curvature and the force record of the owned fluid nodes of plane q
        // (src/main_iteration_GPU.cu:908-1003)
        if (p - 4 >= k0) {
            const int q = p - 4;
            const T* cl[3] = {&s_cn[s4(q - 1)][0][0], &s_cn[s4(q)][0][0], &s_cn[s4(q + 1)][0][0]};
            const T* cno = s_cnorm[s3(q)];
            const int16_t* list = s_l7[par];
            const int n = min(count(par, 3), NO);
            static_assert(NO <= GRAD_THREADS, "one owned node per thread");
            if (tid < n) {
                const int e = pos(list[tid]), x = e % WX, y = e / WX;
                const int v = (x - 2) + VX * (y - 2);
                auto comp = [&](int c) {
                    return [&, c](int dx, int dy, int dz) -> T { return cl[1 + dz][c * NV + v + dx + VX * dy]; };
                };
                // away from interfaces c_norm and the normals are zero and so is the force
                // record, whatever the curvature: a wave with no interface node skips it
                const T c_norm = cno[(x - H) + TX * (y - H)];
                T curv = T(0);
                if (c_norm != T(0)) curv = curvature_of<T>(comp(0), comp(1), comp(2));
                T* a = A.aux + 4 * (int64_t)rec;
                a[0] = cl[1][v];
                a[1] = cl[1][NV + v];
                a[2] = cl[1][2 * NV + v];
                a[3] = T(0.5) * P.gamma * curv * c_norm;
            }
        }
    };
    for (int p = k0 - H; p <= k1 + H; p += 2) {
        body(p, std::integral_constant<int, 0>{});
        if (p + 1 <= k1 + H) body(p + 1, std::integral_constant<int, 1>{});
    }
}


// ---------------------------------------------------------------------------------------
// Two-pass chain on 3-D tiles.  The z-march above advances one plane per pass of its five
// dependent stages (about 5 us a plane at two workgroups per CU whatever the tile area),
// so its time is set by latency where its z-runs are short.  Here the chain is cut where
// the normals are complete: k_grad_normals computes K3 and K4+K5 on a box of nodes and
// writes the normals and c_norm of its fluid nodes to a dense S2 array of 4-vectors
// (Vec4: the four values of a node side by side, one vector store); k_grad_force reads
// them back, extrapolates them to the solid-boundary nodes (K6) and computes the
// curvature and force records (K7).  Each box is loaded once with a halo of two nodes,
// every workgroup runs three short phases, and many small workgroups per CU hide the
// latency.  Node sets and arithmetic are those of the march: K3 over [-2, n+3], K4 over
// [-1, n+2], K6 over [0, n+1], K7 over [1, n], through the same device functions.
//
// A workgroup has one thread per column (x, y) of its box (LX * LY threads); the thread
// walks its column up z, so that its loads, LDS stores and list flags take one address
// step per plane.  Box entry e = x + LX * (y + LY * z).

// Stage lists of a box, built from the tag words each thread loaded: a wave counts its
// entries with ballots and reserves its runs with one LDS atomic on the packed counts.
struct BoxLists {
    template <int PER, class F0, class F1, class W0, class W1>
    __device__ __forceinline__ static void build(unsigned* cnt, int lane, F0&& f0, F1&& f1, W0&& w0, W1&& w1) {
        uint32_t c = 0;
#pragma unroll
        for (int m = 0; m < PER; ++m)
            c += (uint32_t)__popcll(__ballot(f0(m))) | (uint32_t)__popcll(__ballot(f1(m))) << 16;
        uint32_t base = 0;
        if (lane == 0) base = atomicAdd(cnt, c);
        base = (uint32_t)__builtin_amdgcn_readfirstlane((int)base);
        int b0 = (int)(base & 0xFFFF), b1 = (int)(base >> 16);
        auto below = [](uint64_t mk) {
            return (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(mk >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mk, 0));
        };
#pragma unroll
        for (int m = 0; m < PER; ++m) {
            const bool a0 = f0(m), a1 = f1(m);
            const uint64_t m0 = __ballot(a0), m1 = __ballot(a1);
            if (a0) w0(m, b0 + below(m0));
            if (a1) w1(m, b1 + below(m1));
            b0 += __popcll(m0); b1 += __popcll(m1);
        }
    }
};

// Owned box (i0, j0, k0) + [0, TX) x [0, TY) x [0, TZ) of this workgroup: boxes laid out x
// fastest, `tx_n` x `ty_n` per z layer, the first at (org, org, org).
__device__ __forceinline__ void box_origin(int tx_n, int ty_n, int tz_n, int order, int org, int TX, int TY, int TZ,
                                           int& i0, int& j0, int& k0) {
    const int b = xcd_block();
    i0 = org + (b % tx_n) * TX;
    if (order == 0) {
        j0 = org + ((b / tx_n) % ty_n) * TY;
        k0 = org + (b / (tx_n * ty_n)) * TZ;
    } else {
        k0 = org + ((b / tx_n) % tz_n) * TZ;
        j0 = org + (b / (tx_n * tz_n)) * TY;
    }
}

template <typename T>
__device__ __forceinline__ void ws_table(T (*s_ws)[13], int tid) {
    if (tid < 7 * 13) {
        const int a = tid / 13, dd = tid % 13;
        T ws = T(0);
        for (int m = 0; m < a; ++m) ws += W<T>::w1;
        for (int m = 0; m < dd; ++m) ws += W<T>::w2;
        s_ws[a][dd] = ws;
    }
}

// K3 + K4/K5 on the box of owned nodes inside [-1, n+2]^3: phi of the solid-boundary nodes
// of the box and its 1-halo, then normals and c_norm of the fluid nodes of the box.
template <typename T, int TX, int TY, int TZ>
__global__ __launch_bounds__((TX + 4) * (TY + 4)) void k_grad_normals(GradArgs<T> A, StepConst<T> P) {
    constexpr int LX = TX + 4, LY = TY + 4, LZ = TZ + 4, NC = LX * LY, NL = NC * LZ, THREADS = NC;
    constexpr int N3 = (TX + 2) * (TY + 2) * (TZ + 2), N4 = TX * TY * TZ;
    static_assert(NL <= 4096, "list entries keep the box position in 12 bits");
    __shared__ T s_phi[NL];
    __shared__ uint32_t s_l3[N3];  // e | fluid-neighbour mask << 12
    __shared__ uint16_t s_l4[N4];  // e | wet << 12
    __shared__ unsigned s_cnt;     // K3 count | K4 count << 16
    __shared__ T s_ws[7][13];
    const Dims& d = A.d;
    const int tid = threadIdx.x, lane = tid & 63;
    int i0, j0, k0;
    box_origin(A.tiles_x, A.tiles_y, A.tiles_z, A.box_order, -1, TX, TY, TZ, i0, j0, k0);
    // owned extent, clipped to [-1, n+2]
    const int ox = min(TX, d.nx + 3 - i0), oy = min(TY, d.ny + 3 - j0), oz = min(TZ, d.nz + 3 - k0);
    ws_table<T>(s_ws, tid);
    if (tid == 0) s_cnt = 0;

    // this thread's column; entries beyond n+4 load a clamped address and count as solid
    const int x = tid % LX, y = tid / LX;
    const int gi = i0 - 2 + x, gj = j0 - 2 + y;
    const bool col = gi <= d.nx + 4 && gj <= d.ny + 4;
    const int64_t P4 = (int64_t)d.X4 * d.Y4;
    const int64_t c0 = i4(d, min(gi, d.nx + 4), min(gj, d.ny + 4), k0 - 2);
    const int zmax = d.nz + 4 - (k0 - 2);  // last loadable z of the box
    T pv[LZ];
    int32_t tv[LZ];
#pragma unroll
    for (int z = 0; z < LZ; ++z) {
        const int64_t c = c0 + (int64_t)min(z, zmax) * P4;
        pv[z] = A.phi[c];
        tv[z] = A.tag[c];
        if (!col || z > zmax) { pv[z] = T(0); tv[z] = TAG_SOLID; }
    }
    __syncthreads();
#pragma unroll
    for (int z = 0; z < LZ; ++z) s_phi[tid + NC * z] = pv[z];
    const bool c3 = x >= 1 && x < ox + 3 && y >= 1 && y < oy + 3;  // 1-halo column
    const bool c4 = x >= 2 && x < ox + 2 && y >= 2 && y < oy + 2;  // owned column
    BoxLists::build<LZ>(
        &s_cnt, lane, [&](int z) { return c3 && z >= 1 && z < oz + 3 && (tv[z] & 3) == TAG_SOLID_BND; },
        [&](int z) { return c4 && z >= 2 && z < oz + 2 && tag_fluid(tv[z] & 3); },
        [&](int z, int at) { s_l3[at] = (uint32_t)(tid + NC * z) | ((uint32_t)tv[z] >> 2) << 12; },
        [&](int z, int at) { s_l4[at] = (uint16_t)((tid + NC * z) | ((tv[z] & 3) == TAG_WET) << 12); });
    __syncthreads();
    const int n3 = min((int)(s_cnt & 0xFFFF), N3), n4 = min((int)(s_cnt >> 16), N4);

    // the solid normals of the wetting nodes among this thread's K4 entries, issued now so
    // that their latency hides behind K3
    constexpr int M4 = (N4 + THREADS - 1) / THREADS;
    T wn[M4][3];
#pragma unroll
    for (int m = 0; m < M4; ++m) {
        const int i = tid + m * THREADS;
        wn[m][0] = wn[m][1] = wn[m][2] = T(0);
        if (i < n4) {
            const int v = s_l4[i], e = v & 0xFFF;
            if ((v >> 12) & 1) {
                const int c = i4(d, i0 - 2 + e % LX, j0 - 2 + (e / LX) % LY, k0 - 2 + e / NC);
                const Vec4<T> sn = reinterpret_cast<const Vec4<T>*>(A.snp)[c];
                wn[m][0] = sn.x; wn[m][1] = sn.y; wn[m][2] = sn.z;
            }
        }
    }

    // K3 (src/main_iteration_GPU.cu:732-755), in place: reads fluid entries, writes
    // solid-boundary ones; the solid-boundary phi of [1, nx] x [1, ny] x [0, nz+1] owned by
    // this box goes back to memory (the later kernels of the step read it there)
    for (int i = tid; i < n3; i += THREADS) {
        const uint32_t lv = s_l3[i];
        const int e = (int)(lv & 0xFFF);
        const uint32_t mask = lv >> 12;
        T ps = T(0);
#pragma unroll
        for (int q = 1; q < 19; ++q) {
            const bool f = (mask >> (q - 1)) & 1u;
            ps += keep_if(f, s_phi[e + EX[q] + LX * (EY[q] + LY * EZ[q])]) * wq<T>(q);
        }
        const T v = ps / s_ws[__popc(mask & 63u)][__popc(mask >> 6)];
        s_phi[e] = v;
        const int bx = e % LX - 2, by = (e / LX) % LY - 2, bz = e / NC - 2;
        const int qi = i0 + bx, qj = j0 + by, qk = k0 + bz;
        if (bx >= 0 && bx < ox && by >= 0 && by < oy && bz >= 0 && bz < oz && qi >= 1 && qi <= d.nx && qj >= 1 &&
            qj <= d.ny && qk >= 0 && qk <= d.nz + 1)
            A.phi[i4(d, qi, qj, qk)] = v;
    }
    __syncthreads();

    // K4 + K5 (src/main_iteration_GPU.cu:757-878)
#pragma unroll
    for (int m = 0; m < M4; ++m) {
        const int i = tid + m * THREADS;
        if (i >= n4) break;
        const int v = s_l4[i], e = v & 0xFFF;
        const bool wet = (v >> 12) & 1;
        auto a = [&](int dx, int dy, int dz) -> T { return s_phi[e + dx + LX * (dy + LY * dz)]; };
        const T cx = iso4_at<T, 0>(a);
        const T cy = iso4_at<T, 1>(a);
        const T cz = iso4_at<T, 2>(a);
        T vx, vy, vz;
        const T cnorm = interface_normal<T>(
            cx, cy, cz, wet, P.cos_theta, [&](T& nwx, T& nwy, T& nwz) { nwx = wn[m][0]; nwy = wn[m][1]; nwz = wn[m][2]; },
            vx, vy, vz);
        const int c2 = i2(d, i0 - 2 + e % LX, j0 - 2 + (e / LX) % LY, k0 - 2 + e / NC);
        reinterpret_cast<Vec4<T>*>(A.cnp)[c2] = Vec4<T>{vx, vy, vz, cnorm};
    }
}

// K6 + K7 on the box of owned nodes inside [1, n]^3: normals of the solid-boundary nodes of
// the box and its 1-halo from the fluid normals around them, then the curvature and the
// force record of the fluid nodes of the box.
template <typename T, int TX, int TY, int TZ>
__global__ __launch_bounds__((TX + 4) * (TY + 4)) void k_grad_force(GradArgs<T> A, StepConst<T> P) {
    constexpr int LX = TX + 4, LY = TY + 4, LZ = TZ + 4, NC = LX * LY, NL = NC * LZ, THREADS = NC;
    constexpr int N6 = (TX + 2) * (TY + 2) * (TZ + 2), N7 = TX * TY * TZ;
    static_assert(NL <= 4096, "list entries keep the box position in 12 bits");
    __shared__ T s_cn[3][NL];
    __shared__ uint32_t s_l6[N6];  // e | fluid-neighbour mask << 12
    __shared__ uint16_t s_l7[N7];  // e
    __shared__ int32_t s_r7[N7];   // force-record index
    __shared__ unsigned s_cnt;     // K6 count | K7 count << 16
    __shared__ T s_ws[7][13];
    const Dims& d = A.d;
    const int tid = threadIdx.x, lane = tid & 63;
    int i0, j0, k0;
    box_origin(A.tiles_x, A.tiles_y, A.tiles_z, A.box_order, 1, TX, TY, TZ, i0, j0, k0);
    const int ox = min(TX, d.nx + 1 - i0), oy = min(TY, d.ny + 1 - j0), oz = min(TZ, d.nz + 1 - k0);
    ws_table<T>(s_ws, tid);
    if (tid == 0) s_cnt = 0;

    // the fluid normals of the box with its 2-halo (inside [-1, n+2], where K4 wrote them),
    // loaded whatever the node type (only fluid entries are read before K6 writes the
    // solid-boundary ones) so that no load waits for the tag; entries beyond n+2 load a
    // clamped address and count as solid
    const int x = tid % LX, y = tid / LX;
    const int gi = i0 - 2 + x, gj = j0 - 2 + y;
    const bool col = gi <= d.nx + 2 && gj <= d.ny + 2;
    const int64_t P4 = (int64_t)d.X4 * d.Y4, P2 = (int64_t)d.X2 * d.Y2;
    const int ci = min(gi, d.nx + 2), cj = min(gj, d.ny + 2);
    const int64_t t0 = i4(d, ci, cj, k0 - 2), n0 = i2(d, ci, cj, k0 - 2);
    const int zmax = d.nz + 2 - (k0 - 2);
    int32_t tv[LZ];
    T cv[LZ][3];
#pragma unroll
    for (int z = 0; z < LZ; ++z) {
        const int zz = min(z, zmax);
        const int64_t c2 = n0 + (int64_t)zz * P2;
        tv[z] = A.tag[t0 + (int64_t)zz * P4];
        const Vec4<T> v = reinterpret_cast<const Vec4<T>*>(A.cnp)[c2];
        cv[z][0] = v.x; cv[z][1] = v.y; cv[z][2] = v.z;
        if (!col || z > zmax) tv[z] = TAG_SOLID;
    }
    __syncthreads();
#pragma unroll
    for (int z = 0; z < LZ; ++z) {
        s_cn[0][tid + NC * z] = cv[z][0];
        s_cn[1][tid + NC * z] = cv[z][1];
        s_cn[2][tid + NC * z] = cv[z][2];
    }
    const bool c6 = x >= 1 && x < ox + 3 && y >= 1 && y < oy + 3;  // 1-halo column
    const bool c7 = x >= 2 && x < ox + 2 && y >= 2 && y < oy + 2;  // owned column
    BoxLists::build<LZ>(
        &s_cnt, lane, [&](int z) { return c6 && z >= 1 && z < oz + 3 && (tv[z] & 3) == TAG_SOLID_BND; },
        [&](int z) { return c7 && z >= 2 && z < oz + 2 && tag_fluid(tv[z] & 3); },
        [&](int z, int at) { s_l6[at] = (uint32_t)(tid + NC * z) | ((uint32_t)tv[z] >> 2) << 12; },
        [&](int z, int at) { s_l7[at] = (uint16_t)(tid + NC * z); s_r7[at] = tv[z] >> 2; });
    __syncthreads();
    const int n6 = min((int)(s_cnt & 0xFFFF), N6), n7 = min((int)(s_cnt >> 16), N7);

    // c_norm of this thread's K7 entries, issued now so that its latency hides behind K6
    constexpr int M7 = (N7 + THREADS - 1) / THREADS;
    T cnv[M7];
#pragma unroll
    for (int m = 0; m < M7; ++m) {
        const int i = tid + m * THREADS;
        cnv[m] = T(0);
        if (i < n7) {
            const int e = s_l7[i];
            cnv[m] = A.cnp[4 * (int64_t)i2(d, i0 - 2 + e % LX, j0 - 2 + (e / LX) % LY, k0 - 2 + e / NC) + 3];
        }
    }

    // K6 (src/main_iteration_GPU.cu:880-906), in place like K3
    for (int i = tid; i < n6; i += THREADS) {
        const uint32_t lv = s_l6[i];
        const int e = (int)(lv & 0xFFF);
        const uint32_t mask = lv >> 12;
        T sx = T(0), sy = T(0), sz = T(0);
#pragma unroll
        for (int q = 1; q < 19; ++q) {
            const bool f = (mask >> (q - 1)) & 1u;
            const int o = e + EX[q] + LX * (EY[q] + LY * EZ[q]);
            sx += keep_if(f, s_cn[0][o]) * wq<T>(q);
            sy += keep_if(f, s_cn[1][o]) * wq<T>(q);
            sz += keep_if(f, s_cn[2][o]) * wq<T>(q);
        }
        const T ws = s_ws[__popc(mask & 63u)][__popc(mask >> 6)];
        s_cn[0][e] = sx / ws; s_cn[1][e] = sy / ws; s_cn[2][e] = sz / ws;
    }
    __syncthreads();

    // K7 (src/main_iteration_GPU.cu:908-1003): away from interfaces c_norm and the normals
    // are zero and so is the record, whatever the curvature
#pragma unroll
    for (int m = 0; m < M7; ++m) {
        const int i = tid + m * THREADS;
        if (i >= n7) break;
        const int e = s_l7[i];
        const T c_norm = cnv[m];
        auto comp = [&](int c) {
            return [&, c](int dx, int dy, int dz) -> T { return s_cn[c][e + dx + LX * (dy + LY * dz)]; };
        };
        T curv = T(0);
        if (c_norm != T(0)) curv = curvature_of<T>(comp(0), comp(1), comp(2));
        T* a = A.aux + 4 * (int64_t)s_r7[i];
        a[0] = s_cn[0][e];
        a[1] = s_cn[1][e];
        a[2] = s_cn[2][e];
        a[3] = T(0.5) * P.gamma * curv * c_norm;
    }
}

}  // namespace dev
}  // namespace mflbm
