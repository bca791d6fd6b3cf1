// nonenone_amd.hpp — C++17 host wrapper over the C ABI of nonenone_amd.h.
//
// Header-only, no reference types: RAII ownership of contexts, assembly
// handles and device arrays, exceptions instead of return codes, and host-array
// convenience calls for batched flash / phase update / assembly.  Device
// arrays are plain hipMalloc allocations; every call is asynchronous on the
// given stream unless it copies results back to host memory.
//
//   nonenone_amd::Context ctx(desc);                       // nn_context_create
//   nonenone_amd::FlashBatch out = nonenone_amd::flash(ctx, P, T, z);   // nn_flash_batch
//
// Link: -lnonenone_amd -lamdhip64 (include paths: this directory, /opt/rocm/include).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "nonenone_amd.h"

namespace nonenone_amd
{

// A failed call of the C ABI (code = NN_E_*).
class Error : public std::runtime_error
{
public:
    Error(int code, const std::string &what) : std::runtime_error(what), code_(code) {}
    int code() const noexcept { return code_; }

private:
    int code_;
};

inline void check(int rc, const char *call)
{
    if (rc != NN_OK)
        throw Error(rc, std::string(call) + ": " + nn_last_error());
}

inline void checkHip(hipError_t e, const char *call)
{
    if (e != hipSuccess)
        throw Error(NN_E_HIP, std::string(call) + ": " + hipGetErrorString(e));
}

// Owning device array of T.
template <class T>
class DeviceArray
{
public:
    DeviceArray() = default;
    explicit DeviceArray(std::size_t n) : n_(n)
    {
        if (n_ > 0)
            checkHip(hipMalloc(reinterpret_cast<void **>(&p_), n_ * sizeof(T)), "hipMalloc");
    }
    DeviceArray(const DeviceArray &) = delete;
    DeviceArray &operator=(const DeviceArray &) = delete;
    DeviceArray(DeviceArray &&o) noexcept : p_(o.p_), n_(o.n_) { o.p_ = nullptr; o.n_ = 0; }
    DeviceArray &operator=(DeviceArray &&o) noexcept
    {
        std::swap(p_, o.p_);
        std::swap(n_, o.n_);
        return *this;
    }
    ~DeviceArray()
    {
        if (p_)
            (void)hipFree(p_);
    }
    T *get() const noexcept { return p_; }
    std::size_t size() const noexcept { return n_; }
    void upload(const T *host, std::size_t n)
    {
        if (n > n_)
            throw Error(NN_E_INVALID_ARGUMENT, "upload larger than the device array");
        checkHip(hipMemcpy(p_, host, n * sizeof(T), hipMemcpyHostToDevice), "hipMemcpy");
    }
    void upload(const std::vector<T> &host) { upload(host.data(), host.size()); }
    std::vector<T> download(std::size_t n) const
    {
        std::vector<T> host(n);
        if (n > 0)
            checkHip(hipMemcpy(host.data(), p_, n * sizeof(T), hipMemcpyDeviceToHost), "hipMemcpy");
        return host;
    }
    std::vector<T> download() const { return download(n_); }

private:
    T *p_ = nullptr;
    std::size_t n_ = 0;
};

// Owning nn_context (fluid parameters resident on the current device).
class Context
{
public:
    explicit Context(const nn_fluid_desc &fluid, const nn_flash_options *options = nullptr)
        : nc_(fluid.nc)
    {
        check(nn_context_create(&fluid, options, &h_), "nn_context_create");
    }
    Context(const Context &) = delete;
    Context &operator=(const Context &) = delete;
    Context(Context &&o) noexcept : h_(o.h_), nc_(o.nc_) { o.h_ = nullptr; }
    ~Context()
    {
        if (h_)
            nn_context_destroy(h_);
    }
    nn_context *get() const noexcept { return h_; }
    int nc() const noexcept { return nc_; }

    void setKernelTiming(bool enable) { check(nn_context_set_kernel_timing(h_, enable ? 1 : 0), "nn_context_set_kernel_timing"); }
    // (summed milliseconds, launches) of one NN_KERNEL_* kind since timing was enabled.
    std::pair<double, int64_t> kernelTime(int kind) const
    {
        double ms = 0.0;
        int64_t launches = 0;
        check(nn_context_kernel_time(h_, kind, &ms, &launches), "nn_context_kernel_time");
        return {ms, launches};
    }

private:
    nn_context *h_ = nullptr;
    int nc_ = 0;
};

// Host copy of a flash batch (ThreePhaseFlashResult as structure of arrays).
struct FlashBatch
{
    std::vector<uint8_t> presence, converged;
    std::vector<int32_t> status, iterations;
    std::vector<double> beta, saturation, x, Z, rho_m, k_go, k_wo;  // n x 3, n x 3, n x 3 x nc, ...
};

// Device storage of a flash batch and its nn_flash_out view.
class DeviceFlashBatch
{
public:
    DeviceFlashBatch(int64_t n, int nc)
        : n_(n), nc_(nc), presence(n), converged(n), status(n), iterations(n), beta(3 * n),
          saturation(3 * n), x(3 * n * nc), Z(3 * n), rho_m(3 * n), k_go(n * nc), k_wo(n * nc)
    {
    }
    nn_flash_out view() const
    {
        return nn_flash_out{presence.get(), converged.get(), status.get(), iterations.get(),
                            beta.get(), saturation.get(), x.get(), Z.get(), rho_m.get(),
                            k_go.get(), k_wo.get()};
    }
    FlashBatch download() const
    {
        return FlashBatch{presence.download(), converged.download(), status.download(),
                          iterations.download(), beta.download(), saturation.download(),
                          x.download(), Z.download(), rho_m.download(), k_go.download(),
                          k_wo.download()};
    }
    int64_t size() const noexcept { return n_; }

private:
    int64_t n_;
    int nc_;

public:
    DeviceArray<uint8_t> presence, converged;
    DeviceArray<int32_t> status, iterations;
    DeviceArray<double> beta, saturation, x, Z, rho_m, k_go, k_wo;
};

// CubicThreePhaseFlash::flash on host arrays (z is n x nc); synchronous.
inline FlashBatch flash(const Context &ctx, const std::vector<double> &P,
                        const std::vector<double> &T, const std::vector<double> &z)
{
    const int64_t n = int64_t(P.size());
    if (T.size() != P.size() || z.size() != P.size() * std::size_t(ctx.nc()))
        throw Error(NN_E_INVALID_ARGUMENT, "flash: P, T and z sizes disagree");
    DeviceArray<double> dP(n), dT(n), dz(z.size());
    dP.upload(P);
    dT.upload(T);
    dz.upload(z);
    DeviceFlashBatch out(n, ctx.nc());
    const nn_flash_out view = out.view();
    check(nn_flash_batch(ctx.get(), n, dP.get(), dT.get(), dz.get(), &view, nullptr), "nn_flash_batch");
    checkHip(hipDeviceSynchronize(), "hipDeviceSynchronize");
    return out.download();
}

// FullyCompositionalThreePhaseEquilibrium::updatePhaseState on host arrays, in
// place; returns per-cell status.  primary is n x N_p, phaseState n x (3nc+8).
inline std::vector<int32_t> updatePhaseState(const Context &ctx, bool hasWellUnknown,
                                             const std::vector<double> &T,
                                             std::vector<double> &primary,
                                             std::vector<double> &phaseState,
                                             std::vector<uint8_t> *flags = nullptr)
{
    const int64_t n = int64_t(T.size());
    DeviceArray<double> dT(n), dPrim(primary.size()), dPs(phaseState.size());
    DeviceArray<int32_t> dStatus(n);
    DeviceArray<uint8_t> dFlags(n);
    dT.upload(T);
    dPrim.upload(primary);
    dPs.upload(phaseState);
    check(nn_phase_update_batch(ctx.get(), n, hasWellUnknown ? 1 : 0, dT.get(), dPrim.get(),
                                dPs.get(), dStatus.get(), dFlags.get(), nullptr),
          "nn_phase_update_batch");
    checkHip(hipDeviceSynchronize(), "hipDeviceSynchronize");
    primary = dPrim.download();
    phaseState = dPs.download();
    if (flags)
        *flags = dFlags.download();
    return dStatus.download();
}

} // namespace nonenone_amd
