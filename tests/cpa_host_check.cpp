// Host build of this library's CPA phase evaluation (nonenone_amd/csrc/cpa.hpp)
// for the CPU tests (test tooling): reads a raw nn_fluid_desc and states, prints
// "<status> <Z> <lnphi_0> ... <lnphi_nc-1>" per state with 17 significant digits.
//
// Usage: cpa_host_check <fluid_desc.bin> <states.bin>
//   states.bin: int64 n, then P[n], T[n], x[n][nc], role[n] (int32).
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <vector>

#include "../nonenone_amd/csrc/cpa.hpp"
#include "host_params.hpp"

using namespace nn;

template <int NC>
static void run(const FluidParams &fp, int64_t n, const double *P, const double *T,
                const double *x, const int32_t *role)
{
    for (int64_t s = 0; s < n; ++s)
    {
        CpaTemp<NC> ct;
        cpaTemperature<NC>(fp, T[s], ct);
        double z = 0.0, lp[NC] = {};
        Err err;
        cpaPhase<NC, double>(fp, ct, P[s], T[s], x + s * NC, role[s] != kGas, z, lp, err);
        std::printf("%d %.17g", err.code, z);
        for (int c = 0; c < NC; ++c) std::printf(" %.17g", lp[c]);
        std::printf("\n");
    }
}

int main(int argc, char **argv)
{
    if (argc != 3)
        return 2;
    nn_fluid_desc d;
    FILE *f = std::fopen(argv[1], "rb");
    if (!f || std::fread(&d, sizeof d, 1, f) != 1)
        return 3;
    std::fclose(f);
    FILE *g = std::fopen(argv[2], "rb");
    int64_t n = 0;
    if (!g || std::fread(&n, sizeof n, 1, g) != 1 || n < 0)
        return 4;
    std::vector<double> P(n), T(n), x(size_t(n) * d.nc);
    std::vector<int32_t> role(n);
    if (std::fread(P.data(), 8, n, g) != size_t(n) || std::fread(T.data(), 8, n, g) != size_t(n) ||
        std::fread(x.data(), 8, x.size(), g) != x.size() ||
        std::fread(role.data(), 4, n, g) != size_t(n))
        return 5;
    std::fclose(g);
    const FluidParams fp = paramsFrom(d);
    switch (d.nc)
    {
    case 2: run<2>(fp, n, P.data(), T.data(), x.data(), role.data()); return 0;
    case 3: run<3>(fp, n, P.data(), T.data(), x.data(), role.data()); return 0;
    default: return 6;
    }
}
