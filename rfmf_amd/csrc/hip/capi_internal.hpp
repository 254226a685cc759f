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
