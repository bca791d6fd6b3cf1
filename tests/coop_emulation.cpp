// Host emulation of the wavefront-per-state (kModeCoop) flash (test tooling).
//
// The cooperative pass runs one state on all 64 lanes of a wavefront: RR3
// starting guesses, allocation seeds and PR-CPA density-root bracket points /
// bisection trees go to separate lanes and meet in ballots and shuffles.  Here
// 64 host threads play the 64 lanes: each runs Flash<NC, kModeCoop>::flash on the
// same state with its own lane id, and every wave operation (waveBallot,
// waveShfl) is a rendezvous of all 64 threads.  The serial flash (kModePlain,
// the reference's control flow) runs beside it, and one line per state is
// printed: <index> <identical> <coop presence> <serial presence> <coop status>
// <serial status> <coop beta x3> <serial beta x3>, identical = 1 when the whole
// FlashResult and status agree bit for bit.
//
// Usage: coop_emulation <fluid_desc.bin> <states.bin>
//   states.bin: int64 n, then P[n], T[n], z[n][nc].
#define NN_HOST_WAVE_EMULATION 1

#include <hip/hip_runtime.h>

#include <condition_variable>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <thread>
#include <vector>

#include "../nonenone_amd/csrc/flash.hpp"
#include "host_params.hpp"

namespace
{
constexpr int kLanes = 64;
thread_local int tLane = 0;

// All 64 lanes deposit a value, wait for each other, then read.
struct Rendezvous
{
    std::mutex m;
    std::condition_variable cv;
    int arrived = 0;
    unsigned long generation = 0;
    double d[kLanes];
    int i[kLanes];
    bool b[kLanes];

    void wait()
    {
        std::unique_lock<std::mutex> lk(m);
        const unsigned long gen = generation;
        if (++arrived == kLanes)
        {
            arrived = 0;
            ++generation;
            cv.notify_all();
            return;
        }
        cv.wait(lk, [&] { return generation != gen; });
    }
} g;
} // namespace

namespace nn
{
int nnHostWaveLane() { return tLane; }

unsigned long long nnHostWaveBallot(bool pred)
{
    g.b[tLane] = pred;
    g.wait();
    unsigned long long mask = 0;
    for (int l = 0; l < kLanes; ++l)
        if (g.b[l])
            mask |= 1ull << l;
    g.wait();
    return mask;
}

double nnHostWaveShfl(double v, int src)
{
    g.d[tLane] = v;
    g.wait();
    const double r = g.d[src & (kLanes - 1)];
    g.wait();
    return r;
}

int nnHostWaveShfl(int v, int src)
{
    g.i[tLane] = v;
    g.wait();
    const int r = g.i[src & (kLanes - 1)];
    g.wait();
    return r;
}
} // namespace nn

using namespace nn;

template <int Mode, int NC>
static FlashResult<NC> flashState(const FluidParams &fp, const FlashOptions &op, double P,
                                  double T, const double *z, int &status)
{
    Err err;
    Flash<NC, Mode> fl(fp, op, err);
    double zs[NC];
    for (int c = 0; c < NC; ++c) zs[c] = z[c];
    FlashResult<NC> r = fl.flash(P, T, zs);
    status = err.code;
    if (status == kErrNone)
        status = r.converged ? NN_STATUS_OK : NN_STATUS_NOT_CONVERGED;
    return r;
}

template <int NC>
static bool same(const FlashResult<NC> &a, const FlashResult<NC> &b)
{
    bool eq = a.converged == b.converged && a.presence == b.presence;
    eq = eq && std::memcmp(a.beta, b.beta, sizeof a.beta) == 0;
    eq = eq && std::memcmp(a.sat, b.sat, sizeof a.sat) == 0;
    eq = eq && std::memcmp(a.x, b.x, sizeof a.x) == 0;
    eq = eq && std::memcmp(a.Z, b.Z, sizeof a.Z) == 0;
    eq = eq && std::memcmp(a.rho, b.rho, sizeof a.rho) == 0;
    eq = eq && std::memcmp(a.kg, b.kg, sizeof a.kg) == 0;
    eq = eq && std::memcmp(a.kw, b.kw, sizeof a.kw) == 0;
    return eq;
}

template <int NC>
static void run(const FluidParams &fp, const FlashOptions &op, int64_t n, const double *P,
                const double *T, const double *z)
{
    for (int64_t s = 0; s < n; ++s)
    {
        int serialStatus = 0;
        const FlashResult<NC> serial =
            flashState<kModePlain, NC>(fp, op, P[s], T[s], z + s * NC, serialStatus);
        std::vector<FlashResult<NC>> lanes(kLanes);
        std::vector<int> status(kLanes, 0);
        std::vector<std::thread> pool;
        for (int l = 0; l < kLanes; ++l)
            pool.emplace_back([&, l] {
                tLane = l;
                lanes[l] = flashState<kModeCoop, NC>(fp, op, P[s], T[s], z + s * NC, status[l]);
            });
        for (auto &t : pool) t.join();
        const FlashResult<NC> &coop = lanes[0];  // lane 0 stores in the pass
        const bool identical = same(coop, serial) && status[0] == serialStatus;
        std::printf("%lld %d %d %d %d %d %.17g %.17g %.17g %.17g %.17g %.17g\n", (long long)s,
                    identical ? 1 : 0, int(coop.presence), int(serial.presence), status[0],
                    serialStatus, coop.beta[0], coop.beta[1], coop.beta[2], serial.beta[0],
                    serial.beta[1], serial.beta[2]);
        std::fflush(stdout);
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
    FILE *gsf = std::fopen(argv[2], "rb");
    int64_t n = 0;
    if (!gsf || std::fread(&n, sizeof n, 1, gsf) != 1 || n < 0)
        return 4;
    std::vector<double> P(n), T(n), z(size_t(n) * d.nc);
    if (std::fread(P.data(), 8, n, gsf) != size_t(n) || std::fread(T.data(), 8, n, gsf) != size_t(n) ||
        std::fread(z.data(), 8, z.size(), gsf) != z.size())
        return 5;
    std::fclose(gsf);
    const FluidParams fp = paramsFrom(d);
    const FlashOptions op = defaultOptions();
    switch (d.nc)
    {
    case 3: run<3>(fp, op, n, P.data(), T.data(), z.data()); return 0;
    case 6: run<6>(fp, op, n, P.data(), T.data(), z.data()); return 0;
    default: return 6;
    }
}
