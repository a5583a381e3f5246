// Self-check of the traversal's reciprocal (traverse.h: recip) against the
// IEEE division on every 32-bit float (wf_selftest_reciprocal, wf.h).
#include "traverse.h"

#include <cstdint>

namespace wf {
namespace {

// one thread per 2^8 consecutive bit patterns; grid covers all 2^32
__global__ void k_check_recip(unsigned long long* mismatches, uint32_t* first_bits)
{
    const uint32_t base = (blockIdx.x * blockDim.x + threadIdx.x) << 8;
    uint32_t bad = 0, first = 0xFFFFFFFFu;
    for (uint32_t k = 0; k < 256u; ++k) {
        const uint32_t bits = base + k;
        const float x = __uint_as_float(bits);
        const float a = trav::recip(x);
        const float b = 1.0f / x;
        const bool same = __float_as_uint(a) == __float_as_uint(b) || (a != a && b != b);
        if (!same) {
            ++bad;
            if (bits < first) first = bits;
        }
    }
    if (bad) {
        atomicAdd(mismatches, (unsigned long long)bad);
        atomicMin(first_bits, first);
    }
}

} // namespace
} // namespace wf

extern "C" __attribute__((visibility("default"))) wf_status wf_selftest_reciprocal(uint64_t* out_mismatches, uint32_t* out_first_bits)
{
    if (!out_mismatches || !out_first_bits) return WF_ERROR_INVALID_ARGUMENT;
    void* d = nullptr;
    if (hipMalloc(&d, 16) != hipSuccess) return WF_ERROR_OUT_OF_MEMORY;
    unsigned long long* mism = (unsigned long long*)d;
    uint32_t* first = (uint32_t*)((char*)d + 8);
    const uint32_t init[4] = {0u, 0u, 0xFFFFFFFFu, 0u};
    hipError_t e = hipMemcpy(d, init, 16, hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        // 2^32 inputs / 256 per thread = 2^24 threads
        wf::k_check_recip<<<(1u << 24) / 256u, 256>>>(mism, first);
        e = hipGetLastError();
    }
    uint32_t out[4] = {0, 0, 0, 0};
    if (e == hipSuccess) e = hipMemcpy(out, d, 16, hipMemcpyDeviceToHost);
    hipFree(d);
    if (e != hipSuccess) return WF_ERROR_DEVICE;
    *out_mismatches = (uint64_t)out[0] | ((uint64_t)out[1] << 32);
    *out_first_bits = out[2];
    return WF_OK;
}
