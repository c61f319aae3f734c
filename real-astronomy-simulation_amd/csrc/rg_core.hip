// Error reporting, version and the per-device scratch buffer.
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <utility>

#include "rg_internal.h"

namespace {
thread_local char g_err[512] = "";

struct Scratch {
    void *ptr = nullptr;
    size_t bytes = 0;
};
std::mutex g_scratch_mu;
// one buffer per (device, stream): calls on different streams never share one
std::map<std::pair<int, hipStream_t>, Scratch> g_scratch;
}  // namespace

namespace rg {

int g_sort_ticket = 0;
int g_sort_items = 0;
int g_sort_block = 0;
int g_sort_sample = -1;
int g_spin_limit = kDefaultSpinLimit;

void set_error(const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}

int hip_fail(hipError_t e, const char *what, const char *file, int line)
{
    set_error("HIP error %d (%s) in %s at %s:%d", static_cast<int>(e), hipGetErrorString(e),
              what, file, line);
    return RG_ERR_HIP;
}

int device_scratch(size_t bytes, hipStream_t stream, void **ptr)
{
    int dev = 0;
    RG_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    Scratch &s = g_scratch[std::make_pair(dev, stream)];
    if (s.bytes < bytes) {
        // Grow geometrically.  hipFree waits for the device, so a kernel
        // still reading the old buffer finishes first.
        size_t nb = s.bytes ? s.bytes : 4096;
        while (nb < bytes) nb *= 2;
        if (s.ptr) RG_HIP(hipFree(s.ptr));
        s.ptr = nullptr;
        s.bytes = 0;
        RG_HIP(hipMalloc(&s.ptr, nb));
        s.bytes = nb;
    }
    *ptr = s.ptr;
    return RG_OK;
}

}  // namespace rg

extern "C" int rg_version(void) { return RG_ABI_VERSION; }

extern "C" const char *rg_last_error(void) { return g_err; }

extern "C" int rg_set_param(const char *name, int64_t value)
{
    RG_REQUIRE(name, RG_ERR_ARG, "null name");
    if (std::strcmp(name, "direct_ipt") == 0) return rg::rg_direct_set_ipt_impl(value);
    if (std::strcmp(name, "direct_split") == 0) return rg::rg_direct_set_split_impl(value);
    if (std::strcmp(name, "direct_sym") == 0) return rg::rg_direct_set_sym_impl(value);
    if (std::strcmp(name, "fp64_pipe") == 0) return rg::rg_direct_set_fp64_pipe_impl(value);
    if (std::strcmp(name, "sort_ticket") == 0) {
        RG_REQUIRE(value == 0 || value == 1, RG_ERR_ARG, "sort_ticket must be 0 or 1");
        rg::g_sort_ticket = static_cast<int>(value);
        return RG_OK;
    }
    if (std::strcmp(name, "sort_items") == 0) {
        RG_REQUIRE(value == 0 || value == 4 || value == 8 || value == 16, RG_ERR_ARG,
                   "sort_items must be 0, 4, 8 or 16");
        rg::g_sort_items = static_cast<int>(value);
        return RG_OK;
    }
    if (std::strcmp(name, "sort_block") == 0) {
        RG_REQUIRE(value == 0 || value == 256 || value == 1024, RG_ERR_ARG,
                   "sort_block must be 0, 256 or 1024");
        rg::g_sort_block = static_cast<int>(value);
        return RG_OK;
    }
    if (std::strcmp(name, "sort_sample") == 0) {
        RG_REQUIRE(value >= -1 && value <= 1, RG_ERR_ARG,
                   "sort_sample must be -1 (auto), 0 (never) or 1 (whenever possible)");
        rg::g_sort_sample = static_cast<int>(value);
        return RG_OK;
    }
    if (std::strcmp(name, "sort_spin_limit") == 0) {
        RG_REQUIRE(value >= -1 && value <= (int64_t(1) << 30), RG_ERR_ARG,
                   "sort_spin_limit must be -1 (default) or in [0, 2^30]");
        rg::g_spin_limit = value < 0 ? rg::kDefaultSpinLimit : static_cast<int>(value);
        return RG_OK;
    }
    if (std::strcmp(name, "fp32c_d64") == 0) return rg::rg_direct_set_fp32c_d64_impl(value);
    if (std::strcmp(name, "walk_block") == 0) return rg::rg_bh_set_walk_block_impl(value);
    if (std::strcmp(name, "walk_mode") == 0) return rg::rg_bh_set_walk_mode_impl(value);
    if (std::strncmp(name, "chunk_", 6) == 0) return rg::rg_bh_set_chunk_impl(name + 6, value);
    rg::set_error("unknown parameter '%s'", name);
    return RG_ERR_ARG;
}
