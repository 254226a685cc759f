// capi_internal.hpp -- definitions shared by the C ABI translation units.
#pragma once

#include <memory>
#include <string>

#include "engine.hpp"

struct rfmf_ctx {
    std::unique_ptr<rfmf::Context> ctx;
};

namespace rfmf {
// Sets the calling thread's rfmf_last_error() message.
void setLastError(const std::string& msg);
}  // namespace rfmf
