// Host harness of encode_short_rows (mf-lbm-cuda_amd/csrc/layout.h) for
// tests/test_short_rows.py: reads nf, the 32-bit rows nb (18 * nf int32) and a validity byte
// per node from argv[1], encodes every wave, and writes to argv[2] the return value (int32),
// then nbb (NB_WAVE int32 per wave) and nb16 (9 * nf uint32).  Built with g++ (the header's
// device qualifiers defined away); it makes no device call.
#include <cstdio>
#include <vector>
#include "layout.h"

int main(int argc, char** argv) {
    if (argc != 3) return 2;
    FILE* in = std::fopen(argv[1], "rb");
    if (!in) return 3;
    int64_t nf = 0;
    if (std::fread(&nf, sizeof(nf), 1, in) != 1) return 4;
    std::vector<int32_t> nb((size_t)18 * nf);
    std::vector<uint8_t> valid((size_t)nf);
    if (std::fread(nb.data(), sizeof(int32_t), nb.size(), in) != nb.size()) return 5;
    if (std::fread(valid.data(), 1, valid.size(), in) != valid.size()) return 6;
    std::fclose(in);
    const int64_t nw = (nf + 63) / 64;
    std::vector<int32_t> nbb((size_t)nw * mflbm::NB_WAVE, -7);
    std::vector<uint32_t> nb16((size_t)9 * nf, 0xDEADBEEFu);
    const int32_t r = mflbm::encode_short_rows(nb.data(), nf, 0, nw, [&](int64_t t) { return valid[t] != 0; },
                                               nb16.data(), nbb.data());
    FILE* out = std::fopen(argv[2], "wb");
    if (!out) return 7;
    std::fwrite(&r, sizeof(r), 1, out);
    std::fwrite(nbb.data(), sizeof(int32_t), nbb.size(), out);
    std::fwrite(nb16.data(), sizeof(uint32_t), nb16.size(), out);
    std::fclose(out);
    return 0;
}
