// capi_internal.hpp -- definitions shared by the C ABI translation units.
#pragma once

#include <memory>
#include <new>
#include <stdexcept>
#include <string>

#include "../../../include/rfmf.h"
#include "../host/host_common.hpp"
#include "engine.hpp"

struct rfmf_ctx {
    std::unique_ptr<rfmf::Context> ctx;
};

namespace rfmf {
// Sets the calling thread's rfmf_last_error() message.
void setLastError(const std::string& msg);

// One-shot argument checks shared by the annealGpu replacements: a valid CSR
// view, at least one chain with seeds and initial spins, phases and iters
// present when there are phases.  Throws Error(InvalidArg).
void checkOneShot(const rfmf_csr* g, const int8_t* init, int64_t C, const uint64_t* phases, const uint64_t* iters,
                  int64_t nPhases, const uint64_t* seeds);
// n == 0 or no phases: the initial state returned without device work
// (src/anneal_kernels.cu:856-868).  Host memory only.
void passThrough(const rfmf_csr* g, const int8_t* init, int64_t C, int8_t* end, int8_t* best, float* ms,
                 int64_t* bestVals, int64_t* finalVals);

// HBM bytes a job keeps resident with state words of `bits` bits.
int64_t footprintBytes(int64_t n, int64_t m2, int64_t chains, int64_t beta_runs, int64_t n_phases, int bits);

// Runs f and maps what it throws to a status code and the thread's
// rfmf_last_error() message (cleared on success).
template <class F>
rfmf_status guarded(F&& f) {
    try {
        setLastError(std::string());
        f();
        return RFMF_OK;
    } catch (const Error& e) {
        setLastError(e.what());
        return static_cast<rfmf_status>(e.status());
    } catch (const std::bad_alloc&) {
        setLastError("host allocation failed");
        return RFMF_ERR_OUT_OF_MEMORY;
    } catch (const std::exception& e) {
        setLastError(e.what());
        return RFMF_ERR_INVALID_ARG;
    }
}
}  // namespace rfmf
