// ISO 532-1:2017 stationary loudness (Zwicker, frontal free field) per cell from 28 third-octave
// levels, 25 Hz - 12.5 kHz: physics/src/loudness.rs:99-182 (core loudness, the specific loudness
// pattern and its area), the reference card's loud_zwicker (paint-gpu/kernels/loudness.cuh:85-153).
//
// A work-item per cell; levels are read with a cell stride and a band stride, so band-major planes
// (cell stride 1) load coalesced: each of the 28 loads of a wave is 256 contiguous bytes. The
// method's tables are the standard's and sit in constant memory. FP32 throughout; the reference's
// card also runs this part without fast math.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>

#include "../../include/qm_paint.h"
#include "qm_internal.hpp"
#include "qm_zwicker.hpp"

__global__ __launch_bounds__(256) void zwicker_kernel(const float* __restrict__ levels, size_t n, size_t cell_stride,
                                                      size_t band_stride, float* __restrict__ sone) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float lv[28];
#pragma unroll
    for (int b = 0; b < 28; ++b) lv[b] = __builtin_nontemporal_load(levels + i * cell_stride + (size_t)b * band_stride);
    __builtin_nontemporal_store(zwicker_sone(lv), sone + i);
}

namespace {
int zfail(char* err, size_t errlen, const char* fmt, ...) {
    if (err && errlen) {
        va_list args;
        va_start(args, fmt);
        vsnprintf(err, errlen, fmt, args);
        va_end(args);
    }
    return 1;
}
}  // namespace

extern "C" int qm_loudness_device(qm_gpu* gpu, const float* d_levels, size_t n, size_t cell_stride, size_t band_stride,
                                  float* d_sone, void* stream, char* err, size_t errlen) {
    if (!gpu || (n && (!d_levels || !d_sone))) return zfail(err, errlen, "null argument");
    if (n == 0) return 0;
    const size_t blocks = (n + 255) / 256;
    if (blocks > 0x7fffffffu) return zfail(err, errlen, "too many cells");
    const hipError_t d = hipSetDevice(gpu->ordinal);
    if (d != hipSuccess) return zfail(err, errlen, "hipSetDevice: %s", hipGetErrorString(d));
    hipLaunchKernelGGL(zwicker_kernel, dim3((unsigned)blocks), dim3(256), 0, (hipStream_t)stream, d_levels, n,
                       cell_stride, band_stride, d_sone);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return zfail(err, errlen, "zwicker_kernel: %s", hipGetErrorString(e));
    return 0;
}

extern "C" int qm_loudness(qm_gpu* gpu, const float* levels, size_t n, float* sone, char* err, size_t errlen) {
    if (!gpu || (n && (!levels || !sone))) return zfail(err, errlen, "null argument");
    if (n == 0) return 0;
    float *d_levels = nullptr, *d_sone = nullptr;
    hipError_t e = hipSetDevice(gpu->ordinal);
    if (e == hipSuccess) e = hipMalloc((void**)&d_levels, sizeof(float) * 28 * n);
    if (e == hipSuccess) e = hipMalloc((void**)&d_sone, sizeof(float) * n);
    if (e == hipSuccess) e = hipMemcpy(d_levels, levels, sizeof(float) * 28 * n, hipMemcpyHostToDevice);
    int rc = 0;
    if (e == hipSuccess) {
        rc = qm_loudness_device(gpu, d_levels, n, 28, 1, d_sone, nullptr, err, errlen);
        if (!rc) e = hipMemcpy(sone, d_sone, sizeof(float) * n, hipMemcpyDeviceToHost);
    }
    if (d_levels) (void)hipFree(d_levels);
    if (d_sone) (void)hipFree(d_sone);
    if (rc) return rc;
    if (e != hipSuccess) return zfail(err, errlen, "qm_loudness: %s", hipGetErrorString(e));
    return 0;
}
