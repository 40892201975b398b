// pp_copy.hip -- streaming device copy: the achievable-bandwidth yardstick the benchmark
// states beside each stage's HBM rate (16 bytes moved per element, like one FP64 stage per
// grid point).  No counterpart in the reference.
//
// Each thread moves four 16-byte vectors strided by the block width, so a wave issues
// 1 KiB-wide coalesced loads and every block streams 16 KiB; one pass, no loop.  Mode 1 marks
// the accesses nontemporal (streaming, not kept in the caches), mode 2 only the stores, mode 3
// only the loads.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "pittpack_hip.h"

namespace {

constexpr int kThreads = 256, kVec = 4;

template <bool NTL, bool NTS>
__global__ void __launch_bounds__(kThreads) k_copy(const double2 *__restrict__ in, double2 *__restrict__ out, long n2)
{
    const long base = (long)blockIdx.x * kThreads * kVec + threadIdx.x;
    double2 v[kVec];
#pragma unroll
    for (int k = 0; k < kVec; k++) {
        const long i = base + (long)k * kThreads;
        if (i < n2) {
            if constexpr (NTL) {
                v[k].x = __builtin_nontemporal_load(&in[i].x);
                v[k].y = __builtin_nontemporal_load(&in[i].y);
            } else {
                v[k] = in[i];
            }
        }
    }
#pragma unroll
    for (int k = 0; k < kVec; k++) {
        const long i = base + (long)k * kThreads;
        if (i < n2) {
            if constexpr (NTS) {
                __builtin_nontemporal_store(v[k].x, &out[i].x);
                __builtin_nontemporal_store(v[k].y, &out[i].y);
            } else {
                out[i] = v[k];
            }
        }
    }
}

__global__ void k_copy_tail(const double *in, double *out, long first, long n)
{
    const long i = first + threadIdx.x;
    if (i < n) out[i] = in[i];
}

}  // namespace

extern "C" int pp_copy_device(const double *d_in, double *d_out, long long n, int mode, void *stream)
{
    if (!d_in || !d_out || n < 0 || mode < 0 || mode > 3) return PP_INPUT_ARGS;
    if (((uintptr_t)d_in | (uintptr_t)d_out) % 16) return PP_INPUT_ARGS;
    hipStream_t s = (hipStream_t)stream;
    const long n2 = (long)(n / 2);
    const long per_block = (long)kThreads * kVec;
    if (n2 > 0) {
        const dim3 grid((unsigned)((n2 + per_block - 1) / per_block));
        const double2 *i2 = (const double2 *)d_in;
        double2 *o2 = (double2 *)d_out;
        switch (mode) {
        case 1: hipLaunchKernelGGL((k_copy<true, true>), grid, dim3(kThreads), 0, s, i2, o2, n2); break;
        case 2: hipLaunchKernelGGL((k_copy<false, true>), grid, dim3(kThreads), 0, s, i2, o2, n2); break;
        case 3: hipLaunchKernelGGL((k_copy<true, false>), grid, dim3(kThreads), 0, s, i2, o2, n2); break;
        default: hipLaunchKernelGGL((k_copy<false, false>), grid, dim3(kThreads), 0, s, i2, o2, n2); break;
        }
    }
    if (n % 2) hipLaunchKernelGGL(k_copy_tail, dim3(1), dim3(64), 0, s, d_in, d_out, 2 * n2, (long)n);
    return hipGetLastError() == hipSuccess ? PP_SUCCESS : PP_HIP_ERROR;
}
