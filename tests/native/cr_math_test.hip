// The short correctly rounded square root, reciprocal and quotient of
// ddfan_index.h against the compiler's correctly rounded sqrtf and '/' (this
// file is built with -fhip-fp32-correctly-rounded-divide-sqrt, as the library
// is) over their whole operand range: every float x in [2^-20, 2^40] for the
// square root, the reciprocal 1/sqrt(x) and the reciprocal 1/x, and for the
// quotient every divisor of two full
// binades at the range ends and the middle against 4096 dividends spread over
// [2^-40, 2^20] (both signs, and 0), plus 2^26 pseudo-random pairs.  And the
// wave form of the Cephes arcsine against asinf_cephes on every float bit
// pattern (consecutive patterns per wave, so waves run both of its paths and
// the small-argument select inside the fast one).  And the clamp of the DD
// kernels' fetch position, floor(c - 0.5) into [0, N] as one v_med3_f32,
// against fminf(fmaxf()) with the weight it leaves, on every float c for
// table sizes N of 1, 512, 736 and 776.
// Prints the number of inputs whose results differ in any bit.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>

#include "ddfan_index.h"

using namespace ddf;

__device__ __forceinline__ uint32_t hash(uint32_t x) {
    x ^= x >> 16;
    x *= 0x7feb352dU;
    x ^= x >> 15;
    x *= 0x846ca68bU;
    x ^= x >> 16;
    return x;
}

// x runs over the bit patterns [lo, hi) of positive floats.
__global__ void k_sqrt(uint32_t lo, uint32_t hi, unsigned long long* bad) {
    unsigned long long nb = 0;
    for (uint32_t u = lo + blockIdx.x * blockDim.x + threadIdx.x; u < hi; u += gridDim.x * blockDim.x) {
        const float x = __uint_as_float(u);
        const float a = sqrt_rn_mid(x), b = sqrtf(x);
        const float ra = rcp_rn_mid(a), rb = 1.0f / b;
        nb += (__float_as_uint(a) != __float_as_uint(b)) + (__float_as_uint(ra) != __float_as_uint(rb)) +
              (__float_as_uint(rcp_rn_mid(x)) != __float_as_uint(1.0f / x));
    }
    if (nb) atomicAdd(bad, nb);
}

// divisors: bit patterns [lo, hi); dividends: 4096 values spread over the range.
__global__ void k_div_grid(uint32_t lo, uint32_t hi, unsigned long long* bad) {
    unsigned long long nb = 0;
    for (uint32_t u = lo + blockIdx.x * blockDim.x + threadIdx.x; u < hi; u += gridDim.x * blockDim.x) {
        const float d = __uint_as_float(u);
        for (int k = 0; k < 4096; ++k) {
            float n;
            if (k == 0) {
                n = 0.0f;
            } else {
                // 2^-40 .. 2^20 in 2047 log steps with varied mantissas, both signs
                const int e = -40 + (int)(((k >> 1) * 60) / 2048);
                const float m = 1.0f + (float)(hash(k) & 0x7fffff) * 0x1p-23f;
                n = ldexpf(m, e);
                if (n > 0x1p20f) n = 0x1p20f;
                if (k & 1) n = -n;
            }
            nb += __float_as_uint(div_rn_mid(n, d)) != __float_as_uint(n / d);
        }
    }
    if (nb) atomicAdd(bad, nb);
}

__global__ void k_div_rand(uint32_t n, unsigned long long* bad) {
    unsigned long long nb = 0;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const uint32_t h1 = hash(2 * i + 1), h2 = hash(2 * i + 2);
        // divisor exponent in [-20, 39], dividend exponent in [-40, 19]
        const float d = ldexpf(1.0f + (float)(h1 & 0x7fffff) * 0x1p-23f, -20 + (int)((h1 >> 23) % 60));
        float a = ldexpf(1.0f + (float)(h2 & 0x7fffff) * 0x1p-23f, -40 + (int)((h2 >> 23) % 60));
        if (h2 & 0x80000000U) a = -a;
        nb += __float_as_uint(div_rn_mid(a, d)) != __float_as_uint(a / d);
    }
    if (nb) atomicAdd(bad, nb);
}

// every 32-bit pattern: u = base + i for i over the grid-stride range
__global__ void k_asin(unsigned long long* bad) {
    unsigned long long nb = 0;
    const unsigned long long n = 1ull << 32;
    for (unsigned long long i = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (unsigned long long)gridDim.x * blockDim.x) {
        const float x = __uint_as_float((uint32_t)i);
        nb += __float_as_uint(asinf_cephes_wave(x)) != __float_as_uint(asinf_cephes(x));
    }
    if (nb) atomicAdd(bad, nb);
}

// fetch position (texel index and weight) with both clamps, every pattern c
__global__ void k_clamp(int N, unsigned long long* bad) {
    unsigned long long nb = 0;
    for (unsigned long long i = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x;
         i < (1ull << 32); i += (unsigned long long)gridDim.x * blockDim.x) {
        const float cm = __uint_as_float((uint32_t)i) - 0.5f;
        const float fa = fminf(fmaxf(floorf(cm), 0.0f), (float)N);
        const float fb = __builtin_amdgcn_fmed3f(floorf(cm), 0.0f, (float)N);
        const float ta = fminf(fmaxf(cm - fa, 0.0f), 1.0f), tb = fminf(fmaxf(cm - fb, 0.0f), 1.0f);
        nb += (int)fa != (int)fb || __float_as_uint(ta) != __float_as_uint(tb);
    }
    if (nb) atomicAdd(bad, nb);
}

#define CK(x)                                                                   \
    do {                                                                        \
        hipError_t e_ = (x);                                                    \
        if (e_ != hipSuccess) {                                                 \
            printf("HIP error %s at line %d\n", hipGetErrorString(e_), __LINE__); \
            return 2;                                                           \
        }                                                                       \
    } while (0)

int main() {
    unsigned long long* d_bad = nullptr;
    CK(hipMalloc(&d_bad, 5 * sizeof(unsigned long long)));
    CK(hipMemset(d_bad, 0, 5 * sizeof(unsigned long long)));
    const uint32_t lo = __builtin_bit_cast(uint32_t, 0x1p-20f), hi = __builtin_bit_cast(uint32_t, 0x1p40f);
    k_sqrt<<<4096, 256>>>(lo, hi + 1, d_bad);
    CK(hipGetLastError());
    const uint32_t b0 = lo, b1 = __builtin_bit_cast(uint32_t, 0x1p-18f);
    const uint32_t m0 = __builtin_bit_cast(uint32_t, 0x1p10f), m1 = __builtin_bit_cast(uint32_t, 0x1p12f);
    const uint32_t e0 = __builtin_bit_cast(uint32_t, 0x1p38f), e1 = hi;
    k_div_grid<<<4096, 256>>>(b0, b1, d_bad + 1);
    k_div_grid<<<4096, 256>>>(m0, m1, d_bad + 1);
    k_div_grid<<<4096, 256>>>(e0, e1 + 1, d_bad + 1);
    CK(hipGetLastError());
    k_div_rand<<<4096, 256>>>(1u << 26, d_bad + 2);
    CK(hipGetLastError());
    k_asin<<<8192, 256>>>(d_bad + 3);
    CK(hipGetLastError());
    for (int N : {1, 512, 736, 776}) k_clamp<<<8192, 256>>>(N, d_bad + 4);
    CK(hipGetLastError());
    unsigned long long bad[5];
    CK(hipMemcpy(bad, d_bad, sizeof(bad), hipMemcpyDeviceToHost));
    printf("sqrt and reciprocal mismatches: %llu of %llu\n", bad[0], 3ull * (hi + 1 - lo));
    printf("quotient grid mismatches: %llu of %llu\n", bad[1],
           4096ull * ((b1 - b0) + (m1 - m0) + (e1 + 1 - e0)));
    printf("quotient random mismatches: %llu of %llu\n", bad[2], 1ull << 26);
    printf("arcsine wave-form mismatches: %llu of %llu\n", bad[3], 1ull << 32);
    printf("fetch clamp mismatches: %llu of %llu\n", bad[4], 4ull << 32);
    CK(hipFree(d_bad));
    printf("%s\n", bad[0] + bad[1] + bad[2] + bad[3] + bad[4] == 0 ? "cr math ok" : "cr math MISMATCH");
    return 0;
}
