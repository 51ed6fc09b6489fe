#pragma once

// Process-wide plumbing between the facade classes and the mdgpu C ABI.
#include <stdexcept>
#include <string>

#include "mdgpu.h"

namespace mdgpu_facade {

// The context every facade object runs on: device MDGPU_DEVICE (default 0), created on first use and kept for
// the life of the process.
md_ctx* context();

// Error thrown for mdgpu failures that have no counterpart in the reference (HIP errors, bad arguments).
class DeviceError : public std::runtime_error {
public:
    using std::runtime_error::runtime_error;
};

// Status -> exception: the reference's own std::runtime_error cases keep their exact messages.
void check(int status);

}  // namespace mdgpu_facade
