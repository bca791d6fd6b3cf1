// serve.hip — device-resident flash/property service (no counterpart in the
// reference, whose runtime assembles one simulation per process).
//
// Many ensemble members keep their cell state on the GPU (primary variables and
// the phase-state vector, which carries each cell's K-values and phase
// hysteresis) and submit Newton-step requests: the hysteretic phase update
// (stability + flash where the active set changes) followed by the residual and
// BSR Jacobian of the member's grid.  The server admits requests continuously:
// a worker thread starts a pending request as soon as a stream slot is free, its
// member has no step in flight, and its output buffers fit the memory budget;
// admitted requests run concurrently on separate HIP streams, and completions
// free their slot and memory for the next ones.  Requests of one member run in
// submission order, one at a time, so each step's result is the same alone and
// in the mix.  Output buffers come from a size-keyed cache (hipFree would
// synchronise the device and stall the overlap); they belong to the request
// until nn_server_release.
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

namespace
{

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
    int32_t np = 0;
    size_t bytes = 0;
    State state = State::Pending;
    int slot = -1;
    int rc = NN_OK;
    double tSubmit = 0.0, tStart = 0.0, tEnd = 0.0;
    Buffers buf;
    hipEvent_t done = nullptr;
};

size_t align256(size_t b) { return (b + 255) & ~size_t(255); }

} // namespace

struct nn_server
{
    int maxInflight = 1;
    size_t budget = 0;
    size_t inUse = 0, peakInUse = 0;
    int inflight = 0, peakInflight = 0;
    int64_t completed = 0;
    int64_t nextTicket = 1;
    std::vector<hipStream_t> streams;
    std::vector<int> slotBusy;
    std::deque<int64_t> pending;
    std::vector<int64_t> running;
    std::unordered_map<int64_t, Request> reqs;
    std::unordered_set<int64_t> busyMembers;
    std::multimap<size_t, void *> cache;  // released buffers by size
    size_t cached = 0;
    std::mutex mu;
    std::condition_variable cvWork, cvDone;
    std::thread worker;
    bool stopping = false;
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
            return serveFail(NN_E_OUT_OF_MEMORY, "server output buffer allocation failed");
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

    // Launches one admitted request on its slot's stream (caller holds mu).
    int launch(Request &r)
    {
        hipStream_t st = streams[size_t(r.slot)];
        char *p = static_cast<char *>(r.buf.base);
        const size_t res = align256(size_t(r.ncell) * r.np * sizeof(double));
        const size_t bsr = align256(size_t(r.nnzb) * r.np * r.np * sizeof(double));
        const size_t st4 = align256(size_t(r.ncell) * sizeof(int32_t));
        double *dRes = reinterpret_cast<double *>(p);
        double *dBsr = reinterpret_cast<double *>(p + res);
        int32_t *dPuStatus = reinterpret_cast<int32_t *>(p + res + bsr);
        int32_t *dAsmStatus = reinterpret_cast<int32_t *>(p + res + bsr + st4);
        uint8_t *dFlags = reinterpret_cast<uint8_t *>(p + res + bsr + 2 * st4);
        const nn_step_request &q = r.req;
        int rc = nn_phase_update_batch(r.ctx, r.ncell, q.has_well_unknown, q.d_T, q.d_primary,
                                       q.d_phase_state, dPuStatus, dFlags, st);
        if (rc == NN_OK)
            rc = nn_assembly_evaluate(q.assembly, q.dt, q.fugacity_scaling, q.d_primary,
                                      q.d_phase_state, q.d_accum_prev, dRes,
                                      q.want_bsr ? dBsr : nullptr, nullptr, dAsmStatus, st);
        if (rc == NN_OK && hipEventRecord(r.done, st) != hipSuccess)
            rc = serveFail(NN_E_HIP, "hipEventRecord failed");
        return rc;
    }

    void finish(int64_t ticket, Request &r, int rc)
    {
        r.rc = rc;
        r.tEnd = now();
        r.state = State::Done;
        if (r.slot >= 0)
            slotBusy[size_t(r.slot)] = 0;
        inUse -= r.bytes;
        --inflight;
        ++completed;
        busyMembers.erase(r.req.member);
        (void)ticket;
    }

    void loop()
    {
        (void)hipSetDevice(device);
        std::unique_lock<std::mutex> lk(mu);
        for (;;)
        {
            // completions
            for (size_t i = 0; i < running.size();)
            {
                Request &r = reqs[running[i]];
                const hipError_t e = hipEventQuery(r.done);
                if (e == hipErrorNotReady)
                {
                    ++i;
                    continue;
                }
                finish(running[i], r, e == hipSuccess ? NN_OK : NN_E_HIP);
                running[i] = running.back();
                running.pop_back();
                cvDone.notify_all();
            }
            // admissions, in submission order; a request whose buffers do not fit
            // the budget blocks later ones (no starvation of large requests)
            for (auto it = pending.begin(); it != pending.end() && inflight < maxInflight;)
            {
                Request &r = reqs[*it];
                if (busyMembers.count(r.req.member))
                {
                    ++it;
                    continue;
                }
                if (inUse + r.bytes > budget && inflight > 0)
                    break;
                int slot = -1;
                for (size_t s = 0; s < slotBusy.size(); ++s)
                    if (!slotBusy[s])
                    {
                        slot = int(s);
                        break;
                    }
                if (slot < 0)
                    break;
                const int64_t ticket = *it;
                it = pending.erase(it);
                r.slot = slot;
                slotBusy[size_t(slot)] = 1;
                busyMembers.insert(r.req.member);
                inUse += r.bytes;
                peakInUse = std::max(peakInUse, inUse);
                ++inflight;
                peakInflight = std::max(peakInflight, inflight);
                r.tStart = now();
                r.state = State::Running;
                int rc = acquire(r.bytes, r.buf);
                if (rc == NN_OK)
                    rc = launch(r);
                if (rc != NN_OK)
                {
                    finish(ticket, r, rc);
                    cvDone.notify_all();
                    continue;
                }
                running.push_back(ticket);
            }
            if (stopping && pending.empty() && running.empty())
                break;
            if (running.empty())
            {
                cvWork.wait(lk, [&] { return stopping || !pending.empty(); });
            }
            else
            {
                // poll completions at ~20 us; new submissions wake us earlier
                cvWork.wait_for(lk, std::chrono::microseconds(20));
            }
        }
    }
};

extern "C" {

const char *nn_server_last_error(void) { return g_serveError; }

int nn_assembly_info(const nn_assembly *a, nn_context **ctx, int64_t *ncell, int64_t *nnzb,
                     int32_t *np);

int nn_server_create(int32_t max_inflight, int64_t memory_budget_bytes, nn_server **out)
{
    if (!out || max_inflight < 1 || max_inflight > 64 || memory_budget_bytes <= 0)
        return serveFail(NN_E_INVALID_ARGUMENT, "max_inflight in 1..64, positive memory budget");
    *out = nullptr;
    auto *s = new (std::nothrow) nn_server;
    if (!s)
        return serveFail(NN_E_OUT_OF_MEMORY, "host allocation failed");
    s->maxInflight = max_inflight;
    s->budget = size_t(memory_budget_bytes);
    if (hipGetDevice(&s->device) != hipSuccess)
    {
        delete s;
        return serveFail(NN_E_NO_DEVICE, "no HIP device");
    }
    s->streams.resize(size_t(max_inflight), nullptr);
    s->slotBusy.assign(size_t(max_inflight), 0);
    for (auto &st : s->streams)
        if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess)
        {
            for (auto &x : s->streams)
                if (x) (void)hipStreamDestroy(x);
            delete s;
            return serveFail(NN_E_HIP, "stream creation failed");
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
    r.bytes = align256(size_t(r.ncell) * r.np * sizeof(double)) +
              align256(size_t(r.nnzb) * r.np * r.np * sizeof(double)) +
              2 * align256(size_t(r.ncell) * sizeof(int32_t)) + align256(size_t(r.ncell));
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
    const size_t resB = align256(size_t(r.ncell) * r.np * sizeof(double));
    const size_t bsrB = align256(size_t(r.nnzb) * r.np * r.np * sizeof(double));
    const size_t st4 = align256(size_t(r.ncell) * sizeof(int32_t));
    res->rc = r.rc;
    res->submit_s = r.tSubmit;
    res->start_s = r.tStart;
    res->end_s = r.tEnd;
    res->ncell = r.ncell;
    res->nnzb = r.nnzb;
    res->np = r.np;
    res->d_residual = p ? reinterpret_cast<double *>(p) : nullptr;
    res->d_bsr = (p && r.req.want_bsr) ? reinterpret_cast<double *>(p + resB) : nullptr;
    res->d_update_status = p ? reinterpret_cast<int32_t *>(p + resB + bsrB) : nullptr;
    res->d_assembly_status = p ? reinterpret_cast<int32_t *>(p + resB + bsrB + st4) : nullptr;
    res->d_update_flags = p ? reinterpret_cast<uint8_t *>(p + resB + bsrB + 2 * st4) : nullptr;
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
    out->inflight = s->inflight;
    out->peak_inflight = s->peakInflight;
    out->memory_in_use = int64_t(s->inUse);
    out->peak_memory_in_use = int64_t(s->peakInUse);
    out->cached_bytes = int64_t(s->cached);
    return NN_OK;
}

int nn_server_destroy(nn_server *s)
{
    if (!s)
        return NN_OK;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stopping = true;
    }
    s->cvWork.notify_all();
    if (s->worker.joinable())
        s->worker.join();
    for (auto &st : s->streams)
        if (st) (void)hipStreamSynchronize(st);
    for (auto &kv : s->reqs)
    {
        s->giveBack(kv.second.buf);
        if (kv.second.done)
            (void)hipEventDestroy(kv.second.done);
    }
    for (auto &kv : s->cache)
        (void)hipFree(kv.second);
    for (auto &st : s->streams)
        if (st) (void)hipStreamDestroy(st);
    delete s;
    return NN_OK;
}

} // extern "C"
