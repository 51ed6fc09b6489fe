#include "MdgpuRuntime.hpp"

#include <cstdlib>
#include <mutex>

namespace mdgpu_facade {

md_ctx* context() {
    static md_ctx* ctx = nullptr;
    static std::once_flag once;
    std::call_once(once, [] {
        const char* env = std::getenv("MDGPU_DEVICE");
        const int device = env ? std::atoi(env) : 0;
        check(md_ctx_create(device, &ctx));
    });
    return ctx;
}

void check(int status) {
    if (status == MD_OK) return;
    const std::string msg = md_last_error();
    if (status == MD_ERR_RUNTIME) throw std::runtime_error(msg);
    throw DeviceError("mdgpu: " + msg);
}

}  // namespace mdgpu_facade
