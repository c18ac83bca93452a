// Check of the unscaled division (div_rn_unscaled in
// cantucci_amd/csrc/de_device.h) against the compiler's correctly rounded
// `a / b`.
//
// Pass 1, exhaustive: every pair of significands, a and b in [1, 2) --
// 2^46 pairs.  Without over- or underflow every step of both sequences is
// exact under scaling by powers of two, so this covers every pair of normal
// operands whose quotient, reciprocal and remainders stay normal.
//
// Pass 2, random: 2^32 pairs over the range the P8 loop body uses it on,
// b in [2^-60, 2^32] and |a| < 2^39 down to subnormals and zero.  Where a
// and the quotient are at least 2^-100 the two results must be
// bit-identical; always,
// 1 + a/b must round to the same float (the body only uses 1 + a/b).
// (Below 2^-100 the remainder can underflow and the quotient may then be an
// ulp off; those quotients are far below the 2^-25 at which 1 + a/b changes.)
#include <hip/hip_runtime.h>

#include <cstdio>

#include "de_device.h"

__global__ void k_pairs(uint32_t mb_lo, uint32_t mb_hi, unsigned long long* mismatches, unsigned long long* first) {
    const uint32_t ma = blockIdx.x * blockDim.x + threadIdx.x;  // one significand of a per lane
    if (ma >= (1u << 23)) return;
    const float a = __uint_as_float(0x3f800000u | ma);
    unsigned long long bad = 0;
    for (uint32_t mb = mb_lo; mb < mb_hi; ++mb) {
        const float b = __uint_as_float(0x3f800000u | mb);
        const float q0 = cth::div_rn_unscaled(a, b);
        const float q1 = a / b;
        if (__float_as_uint(q0) != __float_as_uint(q1)) {
            ++bad;
            atomicMin(first, ((unsigned long long)ma << 23) | mb);
        }
    }
    if (bad) atomicAdd(mismatches, bad);
}

__device__ __forceinline__ uint64_t mix64(uint64_t z) {
    z += 0x9e3779b97f4a7c15ull;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

__global__ void k_random(uint64_t seed, uint64_t n, unsigned long long* mismatches, unsigned long long* first) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    unsigned long long bad = 0;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const uint64_t h = mix64(seed ^ (i * 0x2545f4914f6cdd1dull));
        // b = 2^eb * 1.mb, eb in [-60, 31]
        const int eb = (int)((h >> 23) % 92u) - 60;
        const float b = __uint_as_float(((uint32_t)(eb + 127) << 23) | (uint32_t)(h & 0x7fffffu));
        // a: sign, exponent field 0..165 (0 = zero or subnormal, 165 -> 2^38), significand
        const uint32_t ea = (uint32_t)((h >> 30) % 166u);
        const uint32_t sa = (uint32_t)(h >> 63) << 31;
        const uint32_t ma = (uint32_t)(h >> 38) & 0x7fffffu;
        const float a = __uint_as_float(sa | (ea << 23) | ((h >> 61) & 1u ? ma : (ma & 0x7ff000u)));
        const float q0 = cth::div_rn_unscaled(a, b);
        const float q1 = a / b;
        // Where a and a/b are both at least 2^-100 the remainders a - b q are
        // multiples of at least 2^-146 and stay exact: bit-identical quotient.
        const bool in_range = fabsf(a) >= 0x1p-100f && fabsf(q1) >= 0x1p-100f;
        const bool ok = (!in_range || __float_as_uint(q0) == __float_as_uint(q1)) &&
                        __float_as_uint(1.0f + q0) == __float_as_uint(1.0f + q1);
        if (!ok) {
            ++bad;
            atomicMin(first, ((unsigned long long)__float_as_uint(a) << 32) | __float_as_uint(b));
        }
    }
    if (bad) atomicAdd(mismatches, bad);
}

static int run_pass(const char* name, unsigned long long* d_mis, unsigned long long* d_first, int which, uint32_t lo,
                    uint32_t hi) {
    unsigned long long zero = 0, init = ~0ull;
    if (hipMemcpy(d_mis, &zero, sizeof(zero), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(d_first, &init, sizeof(init), hipMemcpyHostToDevice) != hipSuccess)
        return 2;
    if (which == 0)
        hipLaunchKernelGGL(k_pairs, dim3((1u << 23) / 256), dim3(256), 0, 0, lo, hi, d_mis, d_first);
    else
        hipLaunchKernelGGL(k_random, dim3(16384), dim3(256), 0, 0, 12345ull, 1ull << 32, d_mis, d_first);
    if (hipDeviceSynchronize() != hipSuccess) {
        std::printf("%s: kernel failed\n", name);
        return 2;
    }
    unsigned long long mis = 0, first = 0;
    if (hipMemcpy(&mis, d_mis, sizeof(mis), hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(&first, d_first, sizeof(first), hipMemcpyDeviceToHost) != hipSuccess)
        return 2;
    std::printf("%s: %llu mismatches", name, mis);
    if (mis) std::printf(" (smallest key 0x%016llx)", first);
    std::printf("\n");
    std::fflush(stdout);
    return mis ? 1 : 0;
}

int main() {
    unsigned long long *d_mis = nullptr, *d_first = nullptr;
    if (hipMalloc(&d_mis, sizeof(unsigned long long)) != hipSuccess ||
        hipMalloc(&d_first, sizeof(unsigned long long)) != hipSuccess) {
        std::printf("hipMalloc failed\n");
        return 2;
    }
    int rc = 0;
    const uint32_t parts = 16, per = (1u << 23) / parts;
    for (uint32_t p = 0; p < parts && rc < 2; ++p) {
        char name[64];
        std::snprintf(name, sizeof name, "significand pairs, b part %u/%u", p + 1, parts);
        rc |= run_pass(name, d_mis, d_first, 0, p * per, (p + 1) * per);
    }
    if (rc < 2) rc |= run_pass("2^32 random pairs in the P8 range", d_mis, d_first, 1, 0, 0);
    return rc;
}
