// Host build of this library's serial phase update (nonenone_amd/csrc/
// phase_update.hpp) for the CPU tests (test tooling): updates every cell in
// kModePlain and writes, per cell, its primary vector, phase-state vector,
// status (int32) and flags (uint8) as raw bytes, so two builds of the same
// source (e.g. with and without NN_CPA_CERTIFIED_SKIP) can be compared bit for bit.
//
// Usage: phase_update_dump <fluid_desc.bin> <cells.bin> <out.bin>
//   cells.bin: int64 n, int64 has_well, then T[n], primary[n][NP], phase_state[n][NPS].
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <vector>

#include "../nonenone_amd/csrc/phase_update.hpp"
#include "host_params.hpp"

using namespace nn;

template <int NC, bool Well>
static void run(const FluidParams &fp, const FlashOptions &op, int64_t n, const double *T,
                const double *prim, const double *ps, FILE *out)
{
    using L = Layout<NC, Well>;
    for (int64_t s = 0; s < n; ++s)
    {
        double pa[L::NP], sa[L::NPS];
        std::memcpy(pa, prim + s * L::NP, sizeof pa);
        std::memcpy(sa, ps + s * L::NPS, sizeof sa);
        int32_t st = -1;
        uint8_t fl = 0;
        phaseUpdateCell<NC, Well, kModePlain>(fp, op, T[s], pa, sa, 0, &st, &fl);
        std::fwrite(pa, sizeof pa, 1, out);
        std::fwrite(sa, sizeof sa, 1, out);
        std::fwrite(&st, sizeof st, 1, out);
        std::fwrite(&fl, sizeof fl, 1, out);
    }
}

static std::vector<char> slurp(const char *path)
{
    std::vector<char> b;
    FILE *f = std::fopen(path, "rb");
    if (!f)
        return b;
    std::fseek(f, 0, SEEK_END);
    b.resize(size_t(std::ftell(f)));
    std::fseek(f, 0, SEEK_SET);
    if (std::fread(b.data(), 1, b.size(), f) != b.size())
        b.clear();
    std::fclose(f);
    return b;
}

int main(int argc, char **argv)
{
    if (argc != 4)
        return 2;
    const std::vector<char> fb = slurp(argv[1]), cb = slurp(argv[2]);
    if (fb.size() != sizeof(nn_fluid_desc) || cb.size() < 16)
        return 3;
    nn_fluid_desc d;
    std::memcpy(&d, fb.data(), sizeof d);
    const FluidParams fp = paramsFrom(d);
    const FlashOptions op = defaultOptions();
    int64_t n, well;
    std::memcpy(&n, cb.data(), 8);
    std::memcpy(&well, cb.data() + 8, 8);
    const int np = 3 * d.nc + 1 + (well ? 1 : 0), nps = 3 * d.nc + 8;
    if (n < 0 || cb.size() != size_t(16 + 8 * n * (1 + np + nps)))
        return 4;
    const double *T = reinterpret_cast<const double *>(cb.data() + 16);
    const double *prim = T + n;
    const double *ps = prim + n * np;
    FILE *out = std::fopen(argv[3], "wb");
    if (!out)
        return 5;
    if (d.nc == 3 && well)
        run<3, true>(fp, op, n, T, prim, ps, out);
    else if (d.nc == 3)
        run<3, false>(fp, op, n, T, prim, ps, out);
    else
    {
        std::fclose(out);
        return 6;
    }
    std::fclose(out);
    return 0;
}
