// serve.hip — device-resident flash/property service (no counterpart in the
// reference, whose runtime assembles one simulation per process).
//
// Many ensemble members keep their cell state on the GPU (primary variables and
// the phase-state vector, which carries each cell's K-values and phase
// hysteresis) and submit Newton-step requests: the hysteretic phase update
// (stability + flash where the active set changes) followed by the residual and
// BSR Jacobian of the member's grid.  A worker thread batches continuously:
// whenever fewer than kMaxBatchesInFlight batches are on the GPU it takes the
// pending steps it can admit (submission order, member idle, output and staging
// buffers within the memory budget, at most maxBatchSteps), gathers the cells of
// steps that share a fluid into one staging array, runs ONE phase-update launch
// over them, scatters the updated states back, and runs each step's assembly.
// A small member alone cannot fill 256 CUs; merged, many can.  Everything goes
// on one stream, in order; per-step events mark completions.  Output and
// staging buffers come from a size-keyed cache (hipFree would synchronise the
// device).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstdio>
#include <deque>
#include <map>
#include <mutex>
#include <new>
#include <thread>
#include <unordered_map>
#include <unordered_set>
#include <vector>

#include "../../include/nonenone_amd.h"
#include "serve_plan.hpp"

namespace
{

using nn::align256;
using nn::GroupPlan;
using nn::stagingBytes;

constexpr int kMaxBatchesInFlight = 2;

thread_local char g_serveError[256] = "";

int serveFail(int code, const char *msg)
{
    std::snprintf(g_serveError, sizeof(g_serveError), "%s", msg);
    return code;
}

struct Buffers
{
    void *base = nullptr;
    size_t bytes = 0;
};

enum class State
{
    Pending,
    Running,
    Done
};

struct Request
{
    nn_step_request req{};
    nn_context *ctx = nullptr;
    int64_t ncell = 0, nnzb = 0;
    int32_t np = 0, nps = 0;
    size_t bytes = 0;  // outputs
    State state = State::Pending;
    int64_t batch = -1;
    int rc = NN_OK;
    double tSubmit = 0.0, tStart = 0.0, tEnd = 0.0;
    Buffers buf;
    hipEvent_t done = nullptr;

    // output layout inside buf: residual, BSR, update status, assembly status, flags
    size_t offBsr() const { return align256(size_t(ncell) * np * sizeof(double)); }
    size_t offStatus() const { return offBsr() + align256(size_t(nnzb) * np * np * sizeof(double)); }
    size_t offAsm() const { return offStatus() + align256(size_t(ncell) * sizeof(int32_t)); }
    size_t offFlags() const { return offAsm() + align256(size_t(ncell) * sizeof(int32_t)); }
    size_t outputBytes() const { return offFlags() + align256(size_t(ncell)); }
};

struct Batch
{
    int64_t id = 0;
    std::vector<int64_t> tickets;
    std::vector<Buffers> staging;
    size_t stagingBytes = 0;
    hipEvent_t done = nullptr;
};

} // namespace

extern "C" int nn_assembly_info(const nn_assembly *a, nn_context **ctx, int64_t *ncell,
                                int64_t *nnzb, int32_t *np);

struct nn_server
{
    int maxBatchSteps = 1;
    size_t budget = 0;
    size_t inUse = 0, peakInUse = 0;
    int inflight = 0, peakInflight = 0;
    int64_t completed = 0, batches = 0, peakBatchSteps = 0, peakMergedCells = 0;
    int64_t nextTicket = 1;
    hipStream_t stream = nullptr;
    std::deque<int64_t> pending;
    std::deque<Batch> running;
    std::unordered_map<int64_t, Request> reqs;
    std::unordered_set<int64_t> busyMembers;
    std::multimap<size_t, void *> cache;  // released buffers by size
    size_t cached = 0;
    std::mutex mu;
    std::condition_variable cvWork, cvDone;
    std::thread worker;
    bool stopping = false;
    bool held = false;  // nn_server_hold: no new batches while set
    int device = 0;
    std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();

    double now() const
    {
        return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }

    // Buffers of at least `bytes` from the cache (smallest that fits) or hipMalloc.
    int acquire(size_t bytes, Buffers &b)
    {
        auto it = cache.lower_bound(bytes);
        if (it != cache.end() && it->first <= bytes + bytes / 4)
        {
            b.base = it->second;
            b.bytes = it->first;
            cached -= it->first;
            cache.erase(it);
            return NN_OK;
        }
        void *p = nullptr;
        if (hipMalloc(&p, bytes) != hipSuccess)
            return serveFail(NN_E_OUT_OF_MEMORY, "server buffer allocation failed");
        b.base = p;
        b.bytes = bytes;
        return NN_OK;
    }

    void giveBack(Buffers &b)
    {
        if (b.base)
        {
            cache.emplace(b.bytes, b.base);
            cached += b.bytes;
        }
        b = Buffers{};
    }

    int gather(const std::vector<int64_t> &grp, int64_t cells, Batch &b, hipStream_t st)
    {
        Request &lead = reqs[grp.front()];
        const int np = lead.np, nps = lead.nps;
        Buffers sb;
        int rc = acquire(stagingBytes(cells, np, nps), sb);
        if (rc != NN_OK)
            return rc;
        b.staging.push_back(sb);
        char *p = static_cast<char *>(sb.base);
        double *sT = reinterpret_cast<double *>(p);
        p += align256(size_t(cells) * sizeof(double));
        double *sP = reinterpret_cast<double *>(p);
        p += align256(size_t(cells) * np * sizeof(double));
        double *sS = reinterpret_cast<double *>(p);
        p += align256(size_t(cells) * nps * sizeof(double));
        int32_t *sSt = reinterpret_cast<int32_t *>(p);
        p += align256(size_t(cells) * sizeof(int32_t));
        uint8_t *sFl = reinterpret_cast<uint8_t *>(p);
        int64_t off = 0;
        for (int64_t t : grp)
        {
            const Request &r = reqs[t];
            if (hipMemcpyAsync(sT + off, r.req.d_T, size_t(r.ncell) * sizeof(double),
                               hipMemcpyDeviceToDevice, st) != hipSuccess ||
                hipMemcpyAsync(sP + off * np, r.req.d_primary, size_t(r.ncell) * np * sizeof(double),
                               hipMemcpyDeviceToDevice, st) != hipSuccess ||
                hipMemcpyAsync(sS + off * nps, r.req.d_phase_state,
                               size_t(r.ncell) * nps * sizeof(double), hipMemcpyDeviceToDevice,
                               st) != hipSuccess)
                return serveFail(NN_E_HIP, "gather copy failed");
            off += r.ncell;
        }
        rc = nn_phase_update_batch(lead.ctx, cells, lead.req.has_well_unknown, sT, sP, sS, sSt, sFl,
                                   st);
        if (rc != NN_OK)
            return rc;
        off = 0;
        for (int64_t t : grp)
        {
            Request &r = reqs[t];
            char *o = static_cast<char *>(r.buf.base);
            if (hipMemcpyAsync(r.req.d_primary, sP + off * np, size_t(r.ncell) * np * sizeof(double),
                               hipMemcpyDeviceToDevice, st) != hipSuccess ||
                hipMemcpyAsync(r.req.d_phase_state, sS + off * nps,
                               size_t(r.ncell) * nps * sizeof(double), hipMemcpyDeviceToDevice,
                               st) != hipSuccess ||
                hipMemcpyAsync(o + r.offStatus(), sSt + off, size_t(r.ncell) * sizeof(int32_t),
                               hipMemcpyDeviceToDevice, st) != hipSuccess ||
                hipMemcpyAsync(o + r.offFlags(), sFl + off, size_t(r.ncell),
                               hipMemcpyDeviceToDevice, st) != hipSuccess)
                return serveFail(NN_E_HIP, "scatter copy failed");
            off += r.ncell;
        }
        return NN_OK;
    }

    // Issues one batch on the stream (caller holds mu): per merge group one
    // phase update (gathered when the group has several steps), then every
    // step's assembly and completion event, then the batch's event.
    int launch(Batch &b, const GroupPlan &plan)
    {
        hipStream_t st = stream;
        int rc = NN_OK;
        for (size_t g = 0; g < plan.groups.size(); ++g)
        {
            const auto &grp = plan.groups[g];
            if (grp.size() > 1)
            {
                peakMergedCells = std::max(peakMergedCells, plan.cells[g]);
                if ((rc = gather(grp, plan.cells[g], b, st)) != NN_OK)
                    return rc;
                continue;
            }
            Request &r = reqs[grp.front()];
            char *o = static_cast<char *>(r.buf.base);
            rc = nn_phase_update_batch(r.ctx, r.ncell, r.req.has_well_unknown, r.req.d_T,
                                       r.req.d_primary, r.req.d_phase_state,
                                       reinterpret_cast<int32_t *>(o + r.offStatus()),
                                       reinterpret_cast<uint8_t *>(o + r.offFlags()), st);
            if (rc != NN_OK)
                return rc;
        }
        for (int64_t t : b.tickets)
        {
            Request &r = reqs[t];
            char *o = static_cast<char *>(r.buf.base);
            const nn_step_request &q = r.req;
            rc = nn_assembly_evaluate(q.assembly, q.dt, q.fugacity_scaling, q.d_primary,
                                      q.d_phase_state, q.d_accum_prev, reinterpret_cast<double *>(o),
                                      q.want_bsr ? reinterpret_cast<double *>(o + r.offBsr())
                                                 : nullptr,
                                      nullptr, reinterpret_cast<int32_t *>(o + r.offAsm()), st);
            if (rc != NN_OK)
                return rc;
            if (hipEventRecord(r.done, st) != hipSuccess)
                return serveFail(NN_E_HIP, "hipEventRecord failed");
        }
        if (hipEventRecord(b.done, st) != hipSuccess)
            return serveFail(NN_E_HIP, "hipEventRecord failed");
        return NN_OK;
    }

    void finish(Request &r, int rc)
    {
        if (r.state == State::Done)
            return;
        r.rc = rc;
        r.tEnd = now();
        r.state = State::Done;
        inUse -= r.bytes;
        --inflight;
        ++completed;
        busyMembers.erase(r.req.member);
    }

    void retire(Batch &b, int rc)
    {
        for (int64_t t : b.tickets) finish(reqs[t], rc);
        for (auto &sb : b.staging) giveBack(sb);
        inUse -= b.stagingBytes;
        if (b.done)
            (void)hipEventDestroy(b.done);
        b.done = nullptr;
    }

    // Forms and issues the next batch; false when nothing could be admitted.
    bool admit()
    {
        Batch b;
        GroupPlan plan;
        std::unordered_set<int64_t> taken;
        size_t outputs = 0;
        for (auto it = pending.begin();
             it != pending.end() && int(b.tickets.size()) < maxBatchSteps;)
        {
            Request &r = reqs[*it];
            if (busyMembers.count(r.req.member) || taken.count(r.req.member))
            {
                ++it;
                continue;
            }
            GroupPlan trial = plan;
            trial.add(*it, r.req.fluid_group, r.ncell, r.np, r.nps, r.req.has_well_unknown);
            const size_t need = outputs + r.bytes + trial.staging();
            if (inUse + need > budget && (inflight > 0 || !b.tickets.empty()))
                break;  // later requests wait: no starvation of large steps
            plan = std::move(trial);
            outputs += r.bytes;
            taken.insert(r.req.member);
            b.tickets.push_back(*it);
            it = pending.erase(it);
        }
        if (b.tickets.empty())
            return false;
        b.id = ++batches;
        b.stagingBytes = plan.staging();
        peakBatchSteps = std::max<int64_t>(peakBatchSteps, int64_t(b.tickets.size()));
        const double t = now();
        int rc = NN_OK;
        for (int64_t tk : b.tickets)
        {
            Request &r = reqs[tk];
            r.state = State::Running;
            r.batch = b.id;
            r.tStart = t;
            busyMembers.insert(r.req.member);
            inUse += r.bytes;
            ++inflight;
            if (rc == NN_OK)
                rc = acquire(r.bytes, r.buf);
        }
        inUse += b.stagingBytes;
        peakInUse = std::max(peakInUse, inUse);
        peakInflight = std::max(peakInflight, inflight);
        if (rc == NN_OK && hipEventCreateWithFlags(&b.done, hipEventDisableTiming) != hipSuccess)
            rc = serveFail(NN_E_HIP, "event creation failed");
        if (rc == NN_OK)
            rc = launch(b, plan);
        if (rc != NN_OK)
        {
            // nothing of this batch is trusted: wait for what was issued, fail every step
            (void)hipStreamSynchronize(stream);
            retire(b, rc);
            cvDone.notify_all();
            return true;
        }
        running.push_back(std::move(b));
        return true;
    }

    void loop()
    {
        (void)hipSetDevice(device);
        std::unique_lock<std::mutex> lk(mu);
        for (;;)
        {
            // per-step completions (steps of a batch finish in issue order)
            bool any = false;
            for (auto &b : running)
                for (int64_t t : b.tickets)
                {
                    Request &r = reqs[t];
                    if (r.state != State::Running)
                        continue;
                    const hipError_t e = hipEventQuery(r.done);
                    if (e == hipErrorNotReady)
                        break;
                    finish(r, e == hipSuccess ? NN_OK : NN_E_HIP);
                    any = true;
                }
            while (!running.empty())
            {
                Batch &b = running.front();
                const hipError_t e = hipEventQuery(b.done);
                if (e == hipErrorNotReady)
                    break;
                retire(b, e == hipSuccess ? NN_OK : NN_E_HIP);
                running.pop_front();
                any = true;
            }
            if (any)
                cvDone.notify_all();
            while (!held && int(running.size()) < kMaxBatchesInFlight && !pending.empty() &&
                   admit())
            {
            }
            if (stopping && pending.empty() && running.empty())
                break;
            if (running.empty())
                cvWork.wait(lk, [&] { return stopping || (!held && !pending.empty()); });
            else
                cvWork.wait_for(lk, std::chrono::microseconds(20));
        }
    }
};

extern "C" {

const char *nn_server_last_error(void) { return g_serveError; }

int nn_server_create(int32_t max_batch_steps, int64_t memory_budget_bytes, void *stream,
                     nn_server **out)
{
    if (!out || max_batch_steps < 1 || max_batch_steps > 4096 || memory_budget_bytes <= 0)
        return serveFail(NN_E_INVALID_ARGUMENT,
                         "max_batch_steps in 1..4096 and a positive memory budget");
    *out = nullptr;
    auto *s = new (std::nothrow) nn_server;
    if (!s)
        return serveFail(NN_E_OUT_OF_MEMORY, "host allocation failed");
    s->maxBatchSteps = max_batch_steps;
    s->budget = size_t(memory_budget_bytes);
    s->stream = static_cast<hipStream_t>(stream);
    if (hipGetDevice(&s->device) != hipSuccess)
    {
        delete s;
        return serveFail(NN_E_NO_DEVICE, "no HIP device");
    }
    s->worker = std::thread([s] { s->loop(); });
    *out = s;
    return NN_OK;
}

int nn_server_submit_step(nn_server *s, const nn_step_request *req, int64_t *ticket)
{
    if (!s || !req || !ticket || !req->assembly || !req->d_T || !req->d_primary ||
        !req->d_phase_state || !req->d_accum_prev)
        return serveFail(NN_E_INVALID_ARGUMENT, "invalid request");
    Request r;
    r.req = *req;
    if (nn_assembly_info(req->assembly, &r.ctx, &r.ncell, &r.nnzb, &r.np) != NN_OK)
        return serveFail(NN_E_INVALID_ARGUMENT, "invalid assembly handle");
    const int nc = (r.np - 1 - (req->has_well_unknown ? 1 : 0)) / 3;
    r.nps = 3 * nc + 8;
    r.bytes = r.outputBytes();
    if (r.bytes > s->budget)
        return serveFail(NN_E_OUT_OF_MEMORY, "request larger than the server's memory budget");
    if (hipEventCreateWithFlags(&r.done, hipEventDisableTiming) != hipSuccess)
        return serveFail(NN_E_HIP, "event creation failed");
    {
        std::lock_guard<std::mutex> lk(s->mu);
        r.tSubmit = s->now();
        const int64_t t = s->nextTicket++;
        s->reqs.emplace(t, r);
        s->pending.push_back(t);
        *ticket = t;
    }
    s->cvWork.notify_one();
    return NN_OK;
}

int nn_server_wait(nn_server *s, int64_t ticket, nn_step_result *res)
{
    if (!s || !res)
        return serveFail(NN_E_INVALID_ARGUMENT, "null argument");
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->reqs.find(ticket);
    if (it == s->reqs.end())
        return serveFail(NN_E_INVALID_ARGUMENT, "unknown ticket");
    // element references survive rehashing by later submissions; iterators do not
    Request &r = it->second;
    s->cvDone.wait(lk, [&] { return r.state == State::Done; });
    char *p = static_cast<char *>(r.buf.base);
    res->rc = r.rc;
    res->np = r.np;
    res->ncell = r.ncell;
    res->nnzb = r.nnzb;
    res->submit_s = r.tSubmit;
    res->start_s = r.tStart;
    res->end_s = r.tEnd;
    res->batch = r.batch;
    res->d_residual = p ? reinterpret_cast<double *>(p) : nullptr;
    res->d_bsr = (p && r.req.want_bsr) ? reinterpret_cast<double *>(p + r.offBsr()) : nullptr;
    res->d_update_status = p ? reinterpret_cast<int32_t *>(p + r.offStatus()) : nullptr;
    res->d_assembly_status = p ? reinterpret_cast<int32_t *>(p + r.offAsm()) : nullptr;
    res->d_update_flags = p ? reinterpret_cast<uint8_t *>(p + r.offFlags()) : nullptr;
    return r.rc;
}

int nn_server_release(nn_server *s, int64_t ticket)
{
    if (!s)
        return serveFail(NN_E_INVALID_ARGUMENT, "null server");
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->reqs.find(ticket);
    if (it == s->reqs.end() || it->second.state != State::Done)
        return serveFail(NN_E_INVALID_ARGUMENT, "unknown or unfinished ticket");
    Request &r = it->second;
    s->giveBack(r.buf);
    if (r.done)
        (void)hipEventDestroy(r.done);
    s->reqs.erase(it);
    return NN_OK;
}

int nn_server_get_stats(nn_server *s, nn_server_stats *out)
{
    if (!s || !out)
        return serveFail(NN_E_INVALID_ARGUMENT, "null argument");
    std::lock_guard<std::mutex> lk(s->mu);
    out->completed = s->completed;
    out->pending = int64_t(s->pending.size());
    out->batches = s->batches;
    out->inflight = s->inflight;
    out->peak_inflight = s->peakInflight;
    out->peak_batch_steps = s->peakBatchSteps;
    out->peak_merged_cells = s->peakMergedCells;
    out->memory_in_use = int64_t(s->inUse);
    out->peak_memory_in_use = int64_t(s->peakInUse);
    out->cached_bytes = int64_t(s->cached);
    return NN_OK;
}

int nn_server_hold(nn_server *s, int32_t hold)
{
    if (!s)
        return serveFail(NN_E_INVALID_ARGUMENT, "null server");
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->held = hold != 0;
    }
    s->cvWork.notify_all();
    return NN_OK;
}

int nn_server_destroy(nn_server *s)
{
    if (!s)
        return NN_OK;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stopping = true;
        s->held = false;  // pending steps still run before the worker ends
    }
    s->cvWork.notify_all();
    if (s->worker.joinable())
        s->worker.join();
    (void)hipStreamSynchronize(s->stream);
    for (auto &kv : s->reqs)
    {
        s->giveBack(kv.second.buf);
        if (kv.second.done)
            (void)hipEventDestroy(kv.second.done);
    }
    for (auto &kv : s->cache)
        (void)hipFree(kv.second);
    delete s;
    return NN_OK;
}

} // extern "C"
