// kt_ctx: one device, one stream, one grow-only device arena.
#pragma once

#include <hip/hip_runtime.h>

#include <mutex>
#include <unordered_map>

#include "../../include/katric.h"
#include "../../include/katric_shard.h"
#include "kt_internal.h"

enum {
    EV_START = 0,
    EV_H2D,
    EV_PREP,
    EV_TC0,
    EV_TC1,
    EV_COUNT,
    EV_END,
    EV_X0,
    EV_X1,
    EV_X2,
    KT_NUM_EVENTS
};

namespace kt {
// Bump allocator over the arena: with base == nullptr it only measures.
struct Carver {
    void* base;
    uint64_t used;
    void* take(uint64_t bytes);
};
struct ServeState;
struct BatchState;
}  // namespace kt

struct kt_ctx {
    int device = 0;
    int num_cus = 0;
    hipStream_t stream = nullptr;
    uint64_t budget = 0;
    void* arena = nullptr;
    uint64_t arena_bytes = 0;
    uint64_t pinned_other = 0;  // device bytes held outside the arena (batch/serve pools)
    hipEvent_t ev[KT_NUM_EVENTS] = {};
    uint64_t* host_small = nullptr;  // pinned staging for flags / totals / stats
    kt::TcParams prm{};
    uint64_t sym_key[2] = {0, 0};  // keys of the symmetry checksum (random per context)
    kt_stats stats{};
    kt_timing timing{};
    std::mutex mu;
    char err[512] = {0};
    kt::ServeState* serve = nullptr;
    kt::BatchState* batch = nullptr;
    std::unordered_map<void*, uint64_t> buffers;  // kt_buffer_alloc allocations
    // two grow-only device buffers outside the arena, for calls whose second
    // stage is sized by a result of the first (the device-side shard
    // producer, triangle listing); their contents last for one call
    void* aux[2] = {nullptr, nullptr};
    uint64_t aux_bytes[2] = {0, 0};

    int fail(int status, const char* fmt, ...);
    int hip_check(hipError_t e, const char* what);
    int reserve(uint64_t bytes);
    int begin_call(bool counting);
    float elapsed(int a, int b);
    // Grows aux[i] to at least `bytes` (its contents are dropped).
    int reserve_aux(int i, uint64_t bytes);
    void release_aux();
    // One device word, read in order after the work queued on `stream`.
    int read_u64(const uint64_t* dev, uint64_t* out, const char* what);
    void destroy_extensions();
};

void kt_batch_release(kt_ctx* c);
void kt_serve_release(kt_ctx* c);

namespace kt {
uint64_t* tc_stats_ptr(uint64_t* ws);
// kt_count_shard without taking the context lock (kt_shard.hip)
int count_shard_locked(kt_ctx* c, const kt_shard_in* in, kt_shard_out* out);
// kt_count without taking the context lock; with fixed_ws the device
// workspace is that region (fixed_ws_bytes long) instead of the context's
// arena, and nothing is allocated (kt_api.hip)
int count_locked(kt_ctx* c, uint32_t n, const uint64_t* offsets, const uint32_t* heads, uint32_t flags,
                 uint64_t* total_out, uint64_t* per_vertex_out, void* fixed_ws, uint64_t fixed_ws_bytes);
// Device address of this device's kt_count_noalloc workspace, looked up once
// per device with the device current (kt_noalloc.hip).
hipError_t noalloc_init(int device);
// KATRIC_TRACE=1: one stderr line per host-side pipeline step
bool trace_enabled();
void trace(const char* fmt, ...);
}
