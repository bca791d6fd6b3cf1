// Host build of the device kernels of one Newton state (test tooling): the
// phase update's passes (lane-per-cell, lockstep stability trials, replay,
// re-flash, cooperative), the accumulation and the residual + BSR assembly, in
// the launch sequence of nn_phase_update_batch / nn_assembly_accumulation /
// nn_assembly_evaluate, with the buffers sized as the library sizes them.  The
// __global__ bodies of nonenone_amd/csrc/kernels.hpp are compiled as host
// functions: each lane-per-item pass runs on one host thread (which drains the
// pass's work queue), the cooperative pass on 64 host threads, one per lane, that
// meet at every wave operation.  Built with -fsanitize=address,undefined it
// checks the kernels' memory accesses on the workloads of the GPU tests.
//
// Usage: kernel_emulation <fluid_desc.bin> <case.bin> <out.bin>
//   case.bin: int64 n, has_well, nconn, nnzb; double dt; then T[n],
//   primary[n][NP], phase_state[n][NPS] (phase-update input), history
//   primary[n][NP] (accumulation input with the same phase state), assembly
//   primary[n][NP], phase_state[n][NPS], previous accumulation[n][NC],
//   volume[n], porosity[n], conn_offsets[n+1], conn_neighbor, conn_trans,
//   conn_gdz (nconn each), row_offsets[n+1], col_blocks[nnzb].
//   out.bin: status[n] (int32), flags[n] (uint8), primary, phase_state,
//   accumulation[n][NC], residual[n][NP], bsr[nnzb][NP][NP], decisions[nconn][3]
//   (uint8), assembly status[n] (int32).
#define NN_HOST_WAVE_EMULATION 1
#include <hip/hip_runtime.h>

#include <atomic>
#include <condition_variable>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <thread>
#include <vector>

struct EmuDim { unsigned x = 0, y = 0, z = 0; };
thread_local EmuDim emuBlockIdx, emuThreadIdx;
EmuDim emuBlockDim;
#undef __global__
#define __global__
#undef __launch_bounds__
#define __launch_bounds__(...)
#undef __device__
#define __device__ __host__
#define blockIdx emuBlockIdx
#define threadIdx emuThreadIdx
#define blockDim emuBlockDim
template <class T> T emuAtomicAdd(T *p, T v) { return __atomic_fetch_add(p, v, __ATOMIC_SEQ_CST); }
#define atomicAdd emuAtomicAdd

#include "../nonenone_amd/csrc/kernels.hpp"
#include "host_params.hpp"

namespace
{
constexpr int kLanes = 64;
thread_local int tLane = 0;
bool gCoop = false;
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
int nnHostWaveLane() { return gCoop ? tLane : 0; }
unsigned long long nnHostWaveBallot(bool pred)
{
    if (!gCoop)
        return pred ? 1ull : 0ull;
    g.b[tLane] = pred;
    g.wait();
    unsigned long long mask = 0;
    for (int l = 0; l < kLanes; ++l)
        if (g.b[l]) mask |= 1ull << l;
    g.wait();
    return mask;
}
double nnHostWaveShfl(double v, int src)
{
    if (!gCoop)
        return v;
    g.d[tLane] = v;
    g.wait();
    const double r = g.d[src & (kLanes - 1)];
    g.wait();
    return r;
}
int nnHostWaveShfl(int v, int src)
{
    if (!gCoop)
        return v;
    g.i[tLane] = v;
    g.wait();
    const int r = g.i[src & (kLanes - 1)];
    g.wait();
    return r;
}
} // namespace nn

using namespace nn;

template <class T>
static std::vector<T> take(FILE *f, size_t n)
{
    std::vector<T> v(n);
    if (n && std::fread(v.data(), sizeof(T), n, f) != n)
    {
        std::fprintf(stderr, "short read\n");
        std::exit(5);
    }
    return v;
}

template <int NC, bool Well>
static int run(const FluidParams &fp, const FlashOptions &op, FILE *in, FILE *out, int64_t n,
               int64_t nconn, int64_t nnzb, double dt)
{
    using L = Layout<NC, Well>;
    auto T = take<double>(in, n);
    auto prim = take<double>(in, n * L::NP);
    auto ps = take<double>(in, n * L::NPS);
    auto histP = take<double>(in, n * L::NP);
    auto rPrim = take<double>(in, n * L::NP);
    auto rPs = take<double>(in, n * L::NPS);
    auto rAcc = take<double>(in, n * NC);
    auto volume = take<double>(in, n);
    auto poro = take<double>(in, n);
    auto co = take<int64_t>(in, n + 1);
    auto nbr = take<int64_t>(in, nconn);
    auto trans = take<double>(in, nconn);
    auto gdz = take<double>(in, nconn);
    auto ro = take<int64_t>(in, n + 1);
    auto cols = take<int64_t>(in, nnzb);
    auto histPs = ps;

    // ---- phase update (nn_phase_update_batch, lockstep on, index order)
    std::vector<int32_t> status(n, -7);
    std::vector<uint8_t> flags(n, 0xEE);
    WorkQueue q{};
    std::vector<unsigned int> list(n), pending(2 * n);
    std::vector<StabRecord<NC>> recs(n);
    emuBlockDim.x = kBlock;
    phaseUpdateKernel<NC, Well>(&fp, &op, n, T.data(), prim.data(), ps.data(), status.data(),
                                flags.data(), &q, list.data(), nullptr, pending.data(),
                                recs.data());
    stabilityTrialKernel<NC>(&fp, &op, &q, pending.data(), recs.data());
    phaseUpdateReplayKernel<NC, Well>(&fp, &op, n, T.data(), prim.data(), ps.data(),
                                      status.data(), flags.data(), &q, list.data(), pending.data(),
                                      recs.data());
    phaseUpdateReflashKernel<NC, Well>(&fp, &op, n, T.data(), prim.data(), ps.data(),
                                       status.data(), flags.data(), &q, list.data(),
                                       pending.data(), recs.data(), nullptr);
    std::fprintf(stderr, "pending %u reflash %u deferred %u\n", q.pending, q.reflash, q.deferred);
    gCoop = true;
    {
        std::vector<std::thread> pool;
        for (int l = 0; l < kLanes; ++l)
            pool.emplace_back([&, l] {
                tLane = l;
                emuThreadIdx.x = unsigned(l);
                phaseUpdateCoopKernel<NC, Well>(&fp, &op, T.data(), prim.data(), ps.data(),
                                                status.data(), flags.data(), &q, list.data());
            });
        for (auto &t : pool) t.join();
    }
    gCoop = false;

    // ---- assembly set-up (nn_assembly_create)
    std::vector<int64_t> self(n), recip(nconn);
    auto slot = [&](int64_t row, int64_t col) -> int64_t {
        for (int64_t b = ro[row]; b < ro[row + 1]; ++b)
            if (cols[b] == col) return b;
        return -1;
    };
    for (int64_t c = 0; c < n; ++c)
    {
        self[c] = slot(c, c);
        for (int64_t k = co[c]; k < co[c + 1]; ++k) recip[k] = slot(nbr[k], c);
    }
    std::vector<double> record(size_t(10 + 3 * NC) * n);
    const bool cpa = fp.model == NN_MODEL_CPA;
    std::vector<double> roots(cpa ? 9 * n : 0);
    AssemblyView v{n, co.data(), nbr.data(), trans.data(), gdz.data(), volume.data(), poro.data(),
                   T.data(), self.data(), recip.data(), record.data(),
                   cpa ? roots.data() : nullptr, nullptr, {}, {}};
    for (int k = 0; k < kMaxNp; ++k) v.rowScale[k] = v.varScale[k] = 1.0;

    // accumulation of the history state (record kernel without record/roots)
    std::vector<double> acc(n * NC);
    {
        AssemblyView va = v;
        va.record = nullptr;
        va.roots = nullptr;
        emuBlockDim.x = 256;
        for (int64_t c = 0; c < n; ++c)
        {
            emuBlockIdx.x = unsigned(c / 256);
            emuThreadIdx.x = unsigned(c % 256);
            cellRecordKernel<NC, Well>(&fp, va, histP.data(), histPs.data(), acc.data(), nullptr);
        }
    }
    // evaluate on the reference's updated state
    std::vector<double> residual(n * L::NP), bsr(size_t(nnzb) * L::NP * L::NP);
    std::vector<uint8_t> dec(nconn * 3);
    std::vector<int32_t> astat(n, -7);
    emuBlockDim.x = 256;
    for (int64_t c = 0; c < n; ++c)
    {
        emuBlockIdx.x = unsigned(c / 256);
        emuThreadIdx.x = unsigned(c % 256);
        cellRecordKernel<NC, Well>(&fp, v, rPrim.data(), rPs.data(), nullptr, astat.data());
    }
    emuBlockDim.x = 64;
    for (int64_t t = 0; t < n * assemblyChunkCount<NC, Well>(); ++t)
    {
        emuBlockIdx.x = unsigned(t / 64);
        emuThreadIdx.x = unsigned(t % 64);
        assembleKernel<NC, Well>(&fp, v, dt, 1.0, rPrim.data(), rPs.data(), rAcc.data(),
                                 residual.data(), bsr.data(), dec.data(), astat.data());
    }
    std::fwrite(status.data(), 4, n, out);
    std::fwrite(flags.data(), 1, n, out);
    std::fwrite(prim.data(), 8, prim.size(), out);
    std::fwrite(ps.data(), 8, ps.size(), out);
    std::fwrite(acc.data(), 8, acc.size(), out);
    std::fwrite(residual.data(), 8, residual.size(), out);
    std::fwrite(bsr.data(), 8, bsr.size(), out);
    std::fwrite(dec.data(), 1, dec.size(), out);
    std::fwrite(astat.data(), 4, n, out);
    return 0;
}

int main(int argc, char **argv)
{
    if (argc != 4)
        return 2;
    nn_fluid_desc d;
    FILE *f = std::fopen(argv[1], "rb");
    if (!f || std::fread(&d, sizeof d, 1, f) != 1)
        return 3;
    std::fclose(f);
    FILE *in = std::fopen(argv[2], "rb");
    FILE *out = std::fopen(argv[3], "wb");
    if (!in || !out)
        return 4;
    int64_t hdr[4];
    double dt;
    if (std::fread(hdr, 8, 4, in) != 4 || std::fread(&dt, 8, 1, in) != 1)
        return 5;
    const FluidParams fp = paramsFrom(d);
    const FlashOptions op = defaultOptions();
    const int64_t n = hdr[0], nconn = hdr[2], nnzb = hdr[3];
    const bool well = hdr[1] != 0;
    int rc = 6;
#define CASE(NCV)                                                                          \
    case NCV:                                                                              \
        rc = well ? run<NCV, true>(fp, op, in, out, n, nconn, nnzb, dt)                    \
                  : run<NCV, false>(fp, op, in, out, n, nconn, nnzb, dt);                  \
        break;
    switch (d.nc)
    {
        CASE(3) CASE(5) CASE(6) CASE(8)
    }
    std::fclose(out);
    return rc;
}
