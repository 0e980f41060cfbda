// Fluid-compact population layout: which (node, slot) cells a step touches and where they
// live.  The boundary kernels' access patterns are replayed here on the host through the
// same code the device runs (zou_he_column, periodic_pdf_copies), so the layout covers
// exactly the cells the reference's dense array would have touched.
#include <hip/hip_runtime.h>
#include <algorithm>
#include <cstdint>
#include <string>
#include <thread>
#include <vector>

#include "layout.h"
#include "lattice.h"
#include "zouhe.h"

namespace mflbm {

bool build_layout(const Dims& d, const std::vector<uint32_t>& fluid, int kper, int jper, int open_z, int in_bc,
                  int out_bc, int plate_z, LayoutHost& out, std::string& err) {
    const int zou_in = open_z && in_bc == 2, zou_out = open_z && out_bc == 2;
    const int64_t S1 = d.S1;
    const int nf = (int)fluid.size();
    out = LayoutHost();
    out.nf = nf;
    std::vector<int32_t> fidx(S1, -1);
    std::vector<int32_t> node(nf);
    for (int t = 0; t < nf; ++t) {
        int i, j, k;
        unpack_ijk(fluid[t], i, j, k);
        node[t] = i1(d, i, j, k);
        fidx[node[t]] = t;
    }
    std::vector<std::vector<int32_t>> ext(19);
    auto note = [&](int64_t n, int s) {
        if (n < 0 || n >= S1) return false;
        if (fidx[n] < 0) ext[s].push_back((int32_t)n);
        return true;
    };
    // 1. odd-step link cells (x + e_q, opp(q))
    for (int t = 0; t < nf; ++t)
        for (int q = 1; q < 19; ++q) {
            const int64_t y = node[t] + EX[q] + (int64_t)d.X1 * (EY[q] + (int64_t)d.Y1 * EZ[q]);
            note(y, OPP[q]);
        }
    // 2. periodic copies of both parities
    std::vector<int32_t> raw[2];
    for (int odd = 0; odd < 2; ++odd)
        periodic_pdf_copies(d, kper, jper, odd != 0, [&](int dst, int src, int s) {
            note(dst, s);
            note(src, s);
            raw[odd].push_back(s);
            raw[odd].push_back(dst);
            raw[odd].push_back(src);
        });
    // 3. Zou-He accesses of both variants
    if (open_z && (zou_in || zou_out)) {
        StepConst<double> P{};
        double dummy = 0.0;
        bool bad = false;
        auto rec = [&](int a, int b, int c, int e, int) -> double& {
            if (!note(i1(d, a, b, c), e)) bad = true;
            return dummy;
        };
        auto wall = [&](int, int, int) -> int { return 0; };
        auto ph = [&](int, int, int) -> double& { return dummy; };
        for (int j = 1; j <= d.ny; ++j)
            for (int i = 1; i <= d.nx; ++i) {
                zou_he_column<double, false>(P, i, j, d.nz, zou_in, zou_out, rec, wall, ph);
                zou_he_column<double, true>(P, i, j, d.nz, zou_in, zou_out, rec, wall, ph);
            }
        if (bad) { err = "Zou-He access outside the population array"; return false; }
    }
    // 4. velocity inlet, convective outlet and porous plate accesses of both variants
    {
        StepConst<double> P{};
        double dummy = 0.0;
        bool bad = false;
        auto rec = [&](int a, int b, int c, int e, int) -> double& {
            if (!note(i1(d, a, b, c), e)) bad = true;
            return dummy;
        };
        auto wall = [&](int, int, int) -> int { return 0; };
        auto ph = [&](int, int, int) -> double& { return dummy; };
        auto cs = [&](int, int) -> double& { return dummy; };
        auto cp = [&]() -> double& { return dummy; };
        for (int j = 1; j <= d.ny; ++j)
            for (int i = 1; i <= d.nx; ++i) {
                if (open_z && in_bc == 1) {
                    velocity_inlet_column<double, false>(P, i, j, 0.0, rec, wall, ph);
                    velocity_inlet_column<double, true>(P, i, j, 0.0, rec, wall, ph);
                }
                if (open_z && out_bc == 1) {
                    convective_outlet_column<double, false>(P, i, j, d.nz, rec, wall, ph, cs, cp);
                    convective_outlet_column<double, true>(P, i, j, d.nz, rec, wall, ph, cs, cp);
                }
                if (plate_z >= 1 && plate_z <= d.nz) {
                    porous_plate_column<false>(i, j, plate_z, 0, rec);
                    porous_plate_column<true>(i, j, plate_z, 0, rec);
                    porous_plate_column<false>(i, j, plate_z, 1, rec);
                    porous_plate_column<true>(i, j, plate_z, 1, rec);
                }
            }
        if (bad) { err = "boundary access outside the population array"; return false; }
    }
    // extra cells: sorted, unique, per slot
    out.ext_off.assign(20, 0);
    size_t mx = 0;
    for (int s = 0; s < 19; ++s) {
        std::sort(ext[s].begin(), ext[s].end());
        ext[s].erase(std::unique(ext[s].begin(), ext[s].end()), ext[s].end());
        out.ext_off[s + 1] = out.ext_off[s] + (int32_t)ext[s].size();
        mx = std::max(mx, ext[s].size());
    }
    out.ns = (int64_t)nf + (int64_t)mx;
    if (out.ns > INT32_MAX) { err = "too many cells for 32-bit cell indices"; return false; }
    out.ext_node.reserve(out.ext_off[19]);
    for (int s = 0; s < 19; ++s) out.ext_node.insert(out.ext_node.end(), ext[s].begin(), ext[s].end());
    auto cell = [&](int64_t n, int s) -> int32_t {
        if (n < 0 || n >= S1) return -1;
        if (fidx[n] >= 0) return fidx[n];
        auto it = std::lower_bound(ext[s].begin(), ext[s].end(), (int32_t)n);
        if (it == ext[s].end() || *it != (int32_t)n) return -1;
        return nf + (int32_t)(it - ext[s].begin());
    };
    // neighbour table
    out.nb.resize((size_t)18 * nf);
    for (int t = 0; t < nf; ++t)
        for (int q = 1; q < 19; ++q) {
            const int64_t y = node[t] + EX[q] + (int64_t)d.X1 * (EY[q] + (int64_t)d.Y1 * EZ[q]);
            const int32_t c = cell(y, OPP[q]);
            if (c < 0 || c >= out.ns) { err = "neighbour cell outside the layout"; return false; }
            out.nb[(size_t)(q - 1) * nf + t] = c;
        }
    // short neighbour rows, encoded over wave ranges on the host's threads
    {
        const int64_t nw = ((int64_t)nf + 63) / 64;
        out.nbb.assign((size_t)nw * NB_WAVE, 0);
        out.nb16.assign((size_t)9 * nf, 0u);
        const unsigned nth = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
        std::vector<int> esc(nth, 0);
        std::vector<std::thread> th;
        for (unsigned k = 0; k < nth; ++k)
            th.emplace_back([&, k] {
                const int64_t a = nw * k / nth, b = nw * (k + 1) / nth;
                esc[k] = encode_short_rows(out.nb.data(), nf, a, b, [](int64_t) { return true; }, out.nb16.data(),
                                           out.nbb.data());
            });
        for (auto& x : th) x.join();
        out.n_escaped = 0;
        for (int e : esc) {
            if (e < 0) { err = "short neighbour row does not decode"; return false; }
            out.n_escaped += e;
        }
    }
    // periodic pairs, last writer wins
    for (int odd = 0; odd < 2; ++odd) {
        const std::vector<int32_t>& r = raw[odd];
        const size_t n = r.size() / 3;
        std::vector<char> keep(n, 1);
        std::vector<std::pair<int64_t, size_t>> key(n);
        for (size_t m = 0; m < n; ++m) key[m] = {(int64_t)r[3 * m] * S1 + r[3 * m + 1], m};
        std::sort(key.begin(), key.end());
        for (size_t m = 0; m + 1 < n; ++m)
            if (key[m].first == key[m + 1].first) keep[key[m].second] = 0;  // a later copy overwrites it
        for (size_t m = 0; m < n; ++m) {
            if (!keep[m]) continue;
            const int s = r[3 * m];
            const int32_t cd = cell(r[3 * m + 1], s), cs = cell(r[3 * m + 2], s);
            if (cd < 0 || cs < 0) { err = "periodic copy of an untracked cell"; return false; }
            out.pairs[odd].push_back(s);
            out.pairs[odd].push_back(cd);
            out.pairs[odd].push_back(cs);
        }
    }
    // porous-plate plane map
    if (plate_z >= 1 && plate_z <= d.nz) {
        const int64_t P2 = (int64_t)d.X1 * d.Y1;
        out.pmap.assign((size_t)3 * 19 * P2, -1);
        for (int p = 0; p < 3; ++p)
            for (int s = 0; s < 19; ++s)
                for (int b = 0; b < d.Y1; ++b)
                    for (int a = 0; a < d.X1; ++a)
                        out.pmap[((size_t)p * 19 + s) * P2 + (size_t)b * d.X1 + a] = cell(i1(d, a, b, plate_z - 1 + p), s);
    }
    // open-boundary plane maps
    if (open_z && (in_bc || out_bc)) {
        const int64_t P2 = (int64_t)d.X1 * d.Y1;
        out.zmap.assign((size_t)6 * 19 * P2, -1);
        const int kk[6] = {0, 1, 2, d.nz - 1, d.nz, d.nz + 1};
        for (int p = 0; p < 6; ++p)
            for (int s = 0; s < 19; ++s)
                for (int b = 0; b < d.Y1; ++b)
                    for (int a = 0; a < d.X1; ++a)
                        out.zmap[((size_t)p * 19 + s) * P2 + (size_t)b * d.X1 + a] = cell(i1(d, a, b, kk[p]), s);
    }
    // every access of the boundary kernels resolves through its plane map, looked up exactly as
    // the kernels do (k_open_bc, k_porous_plate), to a cell of the layout
    {
        const int64_t P2 = (int64_t)d.X1 * d.Y1;
        StepConst<double> P{};
        double dummy = 0.0;
        bool bad = false;
        auto zrec = [&](int a, int b, int c, int e, int) -> double& {
            const int pl = c <= 2 ? c : 3 + (c - (d.nz - 1));
            if (pl < 0 || pl > 5 || a < 0 || a >= d.X1 || b < 0 || b >= d.Y1 || e < 0 || e >= 19) bad = true;
            else {
                const int32_t c2 = out.zmap[((size_t)pl * 19 + e) * P2 + (size_t)b * d.X1 + a];
                if (c2 < 0 || c2 >= out.ns) bad = true;
            }
            return dummy;
        };
        auto prec = [&](int a, int b, int c, int e, int) -> double& {
            const int pl = c - (plate_z - 1);
            if (pl < 0 || pl > 2 || a < 0 || a >= d.X1 || b < 0 || b >= d.Y1 || e < 0 || e >= 19) bad = true;
            else {
                const int32_t c2 = out.pmap[((size_t)pl * 19 + e) * P2 + (size_t)b * d.X1 + a];
                if (c2 < 0 || c2 >= out.ns) bad = true;
            }
            return dummy;
        };
        auto wall = [&](int, int, int) -> int { return 0; };
        auto ph = [&](int, int, int) -> double& { return dummy; };
        auto cs = [&](int, int) -> double& { return dummy; };
        auto cp = [&]() -> double& { return dummy; };
        const bool zmapped = open_z && (in_bc || out_bc);
        for (int j = 1; j <= d.ny && !bad; ++j)
            for (int i = 1; i <= d.nx; ++i) {
                if (zmapped && in_bc == 1) {
                    velocity_inlet_column<double, false>(P, i, j, 0.0, zrec, wall, ph);
                    velocity_inlet_column<double, true>(P, i, j, 0.0, zrec, wall, ph);
                }
                if (zmapped && (zou_in || zou_out)) {
                    zou_he_column<double, false>(P, i, j, d.nz, zou_in, zou_out, zrec, wall, ph);
                    zou_he_column<double, true>(P, i, j, d.nz, zou_in, zou_out, zrec, wall, ph);
                }
                if (zmapped && out_bc == 1) {
                    convective_outlet_column<double, false>(P, i, j, d.nz, zrec, wall, ph, cs, cp);
                    convective_outlet_column<double, true>(P, i, j, d.nz, zrec, wall, ph, cs, cp);
                }
                if (plate_z >= 1 && plate_z <= d.nz)
                    for (int blocked = 0; blocked < 2; ++blocked) {
                        porous_plate_column<false>(i, j, plate_z, blocked, prec);
                        porous_plate_column<true>(i, j, plate_z, blocked, prec);
                    }
            }
        if (bad) { err = "boundary access without a cell in its plane map"; return false; }
    }
    // slice offsets of the fluid list
    out.zstart.assign(d.nz + 1, 0);
    {
        int t = 0;
        for (int k = 1; k <= d.nz; ++k) {
            out.zstart[k - 1] = t;
            while (t < nf) {
                int i, j, kq;
                unpack_ijk(fluid[t], i, j, kq);
                if (kq != k) break;
                ++t;
            }
        }
        out.zstart[d.nz] = nf;
    }
    return true;
}

}  // namespace mflbm
