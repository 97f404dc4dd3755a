// Exhaustive check of axr::rcp_rn (axiomr_amd/csrc/axr_math.hpp) on the GPU:
// for every 32-bit pattern x, rcp_rn(x) must have the bits of the IEEE
// division 1.0f / x (two NaNs count as equal).  Prints the number of inputs
// tested and of mismatches; exit status 0 only if there are none.  Built by
// axiomr_amd/build.py with the library's compiler flags.
#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstdio>

#include "axr_math.hpp"

__global__ void check(uint32_t chunk, unsigned long long* bad, uint32_t* samples, unsigned long long* tested) {
    const uint32_t base = chunk << 28;   // this launch: patterns [chunk << 28, (chunk + 1) << 28)
    unsigned long long n = 0;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < (1u << 28); i += gridDim.x * blockDim.x) {
        const float x = __uint_as_float(base | i);
        const float want = 1.0f / x;
        const float got = axr::rcp_rn(x);
        ++n;
        const bool same = __float_as_uint(want) == __float_as_uint(got) || (want != want && got != got);
        if (!same) {
            const unsigned long long k = atomicAdd(bad, 1ull);
            if (k < 16) samples[k] = base | i;
        }
    }
    atomicAdd(tested, n);
}

int main() {
    unsigned long long *bad = nullptr, *tested = nullptr;
    uint32_t* samples = nullptr;
    if (hipMalloc((void**)&bad, 8) != hipSuccess || hipMalloc((void**)&tested, 8) != hipSuccess ||
        hipMalloc((void**)&samples, 64) != hipSuccess) {
        std::printf("no device memory\n");
        return 2;
    }
    if (hipMemset(bad, 0, 8) != hipSuccess || hipMemset(tested, 0, 8) != hipSuccess ||
        hipMemset(samples, 0, 64) != hipSuccess)
        return 2;
    for (uint32_t chunk = 0; chunk < 16; ++chunk) {
        hipLaunchKernelGGL(check, dim3(4096), dim3(256), 0, 0, chunk, bad, samples, tested);
        if (hipDeviceSynchronize() != hipSuccess) {
            std::printf("chunk %u failed\n", chunk);
            return 2;
        }
    }
    unsigned long long nbad = 0, ntested = 0;
    uint32_t s[16] = {};
    if (hipMemcpy(&nbad, bad, 8, hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(&ntested, tested, 8, hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(s, samples, 64, hipMemcpyDeviceToHost) != hipSuccess)
        return 2;
    std::printf("tested %llu mismatches %llu\n", ntested, nbad);
    for (unsigned long long i = 0; i < nbad && i < 16; ++i) std::printf("  x = 0x%08x\n", s[i]);
    return nbad ? 1 : 0;
}
