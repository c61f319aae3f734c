// Internal helpers shared by the librgsim translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "rgsim.h"

namespace rg {

// Records a printf-style message as this thread's last error.
void set_error(const char *fmt, ...);

// Records a HIP failure and returns RG_ERR_HIP.
int hip_fail(hipError_t e, const char *what, const char *file, int line);

// Grow-only library scratch, one buffer per (current device, stream) (see
// the header's scratch note).  Returns RG_OK and sets *ptr, or an error
// status.
int device_scratch(size_t bytes, hipStream_t stream, void **ptr);

// j-slices of an FP32/FP32C pair sum over n particles with ipt i-particles
// per lane: sets *jlen (a multiple of the 256-particle tile) and returns the
// slice count, ceil(n / *jlen).  Used by the direct sum and its adjoint.
int direct_slices(int n, int ipt, int *jlen);

// Radix sort passes always take tile tickets (1) or only when their tiles
// may not all be resident at once (0, the default).
extern int g_sort_ticket;

// Keys per thread in a radix-sort tile for handles created from now on: 0
// (by n), 4, 8 or 16.
extern int g_sort_items;

// Threads per radix-sort tile for handles created from now on: 0 (by n),
// 256 or 1024 (with 4 keys per thread).
extern int g_sort_block;

// BH tree sort: -1 auto (the sample sort once a handle has an order from
// its previous compute, for 4,096 <= n <= kSampleMaxN), 0 always the LSD
// radix sort, 1 the sample sort whenever a previous order exists.
extern int g_sort_sample;

// Re-polls a decoupled look-back wait (radix sort, treelet ranking) makes
// before it gives up and flags the failure; 0 gives up at once.
extern int g_spin_limit;
constexpr int kDefaultSpinLimit = 1 << 20;

// Tunables behind rg_set_param.
int rg_direct_set_ipt_impl(int64_t v);
int rg_direct_set_split_impl(int64_t v);
int rg_direct_set_sym_impl(int64_t v);
int rg_direct_set_fp64_pipe_impl(int64_t v);
int rg_direct_set_fp32c_d64_impl(int64_t v);
int rg_bh_set_walk_block_impl(int64_t v);
int rg_bh_set_walk_mode_impl(int64_t v);
int rg_bh_set_chunk_impl(const char *what, int64_t v);

inline hipStream_t as_stream(void *s) { return reinterpret_cast<hipStream_t>(s); }

}  // namespace rg

#define RG_HIP(expr)                                                      \
    do {                                                                  \
        hipError_t rg_e_ = (expr);                                        \
        if (rg_e_ != hipSuccess)                                          \
            return rg::hip_fail(rg_e_, #expr, __FILE__, __LINE__);        \
    } while (0)

#define RG_CHECK_LAUNCH() RG_HIP(hipGetLastError())

#define RG_REQUIRE(cond, code, ...)                                       \
    do {                                                                  \
        if (!(cond)) {                                                    \
            rg::set_error(__VA_ARGS__);                                   \
            return (code);                                                \
        }                                                                 \
    } while (0)

// Byte ranges [a, a+na) and [b, b+nb) overlap.
inline bool rg_overlap(const void *a, size_t na, const void *b, size_t nb)
{
    const char *pa = static_cast<const char *>(a);
    const char *pb = static_cast<const char *>(b);
    return pa < pb + nb && pb < pa + na;
}
