// Fill kernels: used in place of hipMemsetAsync so that every pipeline of
// the library stays capturable in a HIP graph.
#include "kt_device.h"
#include "kt_internal.h"

namespace kt {

__global__ void k_fill_u64(uint64_t* __restrict__ p, uint64_t count, uint64_t value) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count;
         i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = value;
}

__global__ void k_fill_u32(uint32_t* __restrict__ p, uint64_t count, uint32_t value) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count;
         i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = value;
}

static unsigned fill_grid(uint64_t count) {
    uint64_t g = (count + 255) / 256;
    if (g > 8192) g = 8192;
    if (g == 0) g = 1;
    return (unsigned)g;
}

void fill_u64(uint64_t* p, uint64_t count, uint64_t value, hipStream_t st) {
    if (count == 0) return;
    k_fill_u64<<<fill_grid(count), 256, 0, st>>>(p, count, value);
}

void fill_u32(uint32_t* p, uint64_t count, uint32_t value, hipStream_t st) {
    if (count == 0) return;
    k_fill_u32<<<fill_grid(count), 256, 0, st>>>(p, count, value);
}

}  // namespace kt
