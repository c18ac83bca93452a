// Exhaustive check of the wave-uniform square root (sqrt_rn_wave in
// cantucci_amd/csrc/de_device.h) against the compiler's correctly rounded
// __builtin_sqrtf, over every one of the 2^32 f32 bit patterns.  Neighbouring
// lanes take neighbouring patterns, so waves mix the fast and the full path
// only at range boundaries; a second pass interleaves patterns from the two
// ends of the range so that every wave also runs the full path.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstring>

#include "de_device.h"

__global__ void k_check(unsigned long long* mismatches, unsigned* first_bad, int mixed) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < (1ull << 32); i += stride) {
        uint32_t bits = (uint32_t)i;
        if (mixed && (i & 1u)) bits ^= 0xfffffffeu;  // odd patterns from the other end of the range
        const float x = __uint_as_float(bits);
        const float a = cth::sqrt_rn_wave(x);
        const float b = __builtin_sqrtf(x);
        const bool same = (__float_as_uint(a) == __float_as_uint(b)) || (a != a && b != b);
        if (!same) {
            atomicAdd(mismatches, 1ull);
            atomicMin(first_bad, bits);
        }
    }
}

int main() {
    unsigned long long* d_mis = nullptr;
    unsigned* d_first = nullptr;
    if (hipMalloc(&d_mis, sizeof(unsigned long long)) != hipSuccess || hipMalloc(&d_first, sizeof(unsigned)) != hipSuccess) {
        std::printf("hipMalloc failed\n");
        return 2;
    }
    int bad = 0;
    for (int mixed = 0; mixed < 2; ++mixed) {
        unsigned long long mis = 0;
        unsigned first = 0xffffffffu;
        if (hipMemcpy(d_mis, &mis, sizeof(mis), hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(d_first, &first, sizeof(first), hipMemcpyHostToDevice) != hipSuccess)
            return 2;
        hipLaunchKernelGGL(k_check, dim3(8192), dim3(256), 0, 0, d_mis, d_first, mixed);
        if (hipDeviceSynchronize() != hipSuccess) {
            std::printf("kernel failed\n");
            return 2;
        }
        if (hipMemcpy(&mis, d_mis, sizeof(mis), hipMemcpyDeviceToHost) != hipSuccess ||
            hipMemcpy(&first, d_first, sizeof(first), hipMemcpyDeviceToHost) != hipSuccess)
            return 2;
        std::printf("pass %d: %llu of 2^32 patterns differ%s", mixed, mis, mis ? "" : "\n");
        if (mis) std::printf(" (first 0x%08x)\n", first);
        bad |= mis != 0;
    }
    return bad ? 1 : 0;
}
