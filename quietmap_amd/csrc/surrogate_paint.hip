// The surrogate painter and its training samples (no counterpart in the reference; include/
// qm_surrogate.h states the contract).
//
//   samples_kernel   - per pair: the surrogate's inputs (calm and favourable rows) and the exact
//                      path's boundary terms per band as targets, with the envelope flag
//   count_kernel     - per receiver: sources within the cutoff (the bucket grid around it)
//   list_kernel      - per receiver: those sources, at the receiver's offset in the pair list
//   features_kernel  - per pair: BF16 input rows for the network, the envelope flag and the source
//                      altitude; pairs outside the envelope are appended to the exact list
//   exact_kernel     - the exact path of every listed pair: its energy per period
//   combine_kernel   - per receiver, in pair order: the network's A_ground + A_dif per band and
//                      state mixed into received energy as the exact path mixes it
//                      (physics/src/ray.rs:155-190, popup/src/evaluate.rs:119-203), or the exact
//                      pair's energy; FP64 sums per receiver.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <new>
#include <vector>

#include "../../include/qm_surrogate.h"
#include "qm_features.hpp"
#include "qm_internal.hpp"

using namespace qm;

namespace {

__device__ __forceinline__ uint16_t to_bf16(float f) {
    uint32_t u = __float_as_uint(f);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

// The bucket range of sources that can lie within `cutoff` of (x, y).
struct BucketBox {
    int kx0, kx1, ky0, ky1;
};
__device__ BucketBox bucket_box(const SourceGrid& sg, double x, double y, double cutoff) {
    BucketBox b;
    b.kx0 = max((int)floor((x - cutoff - sg.origin_x) / sg.bucket_m), 0);
    b.kx1 = min((int)floor((x + cutoff - sg.origin_x) / sg.bucket_m), (int)sg.buckets_x - 1);
    b.ky0 = max((int)floor((y - cutoff - sg.origin_y) / sg.bucket_m), 0);
    b.ky1 = min((int)floor((y + cutoff - sg.origin_y) / sg.bucket_m), (int)sg.buckets_y - 1);
    return b;
}

}  // namespace

__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void samples_kernel(
    Scene sc, BandTable B, const qm_receiver* receivers, const qm_source* sources, const uint32_t* pair_receiver,
    const uint32_t* pair_source, uint32_t n, float* inputs, float* targets, uint32_t* in_envelope) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    PathScratch w;
    StateStream states[2];
    w.state = states;
    const qm_receiver& rcv = receivers[pair_receiver[i]];
    const qm_source src = sources[pair_source[i]];
    float f[FEATURES];
    const SurrogatePath sp = surrogate_features(sc, rcv, src, w.cross, f);
    PathResult r;
    evaluate_path(sc, B, rcv, src, w, r);
    float* x0 = inputs + (size_t)i * 2 * FEATURES;
    for (int k = 0; k < FEATURES; ++k) {
        x0[k] = f[k];
        x0[FEATURES + k] = k == 37 ? 1.0f : f[k];
    }
    float* t0 = targets + (size_t)i * 2 * BANDS;
    for (int b = 0; b < BANDS; ++b) {
        t0[b] = r.att[0][b];
        t0[BANDS + b] = r.att[1][b];
    }
    in_envelope[i] = (sp.in_envelope && r.failed == 0) ? 1u : 0u;
}

__global__ __launch_bounds__(256) void count_kernel(const qm_receiver* receivers, uint32_t n, SourceGrid sg,
                                                    double cutoff, uint32_t* counts) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double rx = receivers[i].x_m, ry = receivers[i].y_m;
    const BucketBox b = bucket_box(sg, rx, ry, cutoff);
    const double cut2 = cutoff * cutoff;
    uint32_t count = 0;
    for (int ky = b.ky0; ky <= b.ky1; ++ky) {
        const uint32_t s0 = sg.bucket_first[(size_t)ky * sg.buckets_x + b.kx0];
        const uint32_t s1 = sg.bucket_first[(size_t)ky * sg.buckets_x + b.kx1 + 1];
        for (uint32_t s = s0; s < s1; ++s) {
            const double dx = sg.sources[s].x_m - rx, dy = sg.sources[s].y_m - ry;
            count += dx * dx + dy * dy <= cut2 ? 1u : 0u;
        }
    }
    counts[i] = count;
}

__global__ __launch_bounds__(256) void list_kernel(const qm_receiver* receivers, uint32_t first, uint32_t n,
                                                   SourceGrid sg, double cutoff, const uint32_t* offsets,
                                                   uint32_t* pair_receiver, uint32_t* pair_source) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t r = first + i;
    const double rx = receivers[r].x_m, ry = receivers[r].y_m;
    const BucketBox b = bucket_box(sg, rx, ry, cutoff);
    const double cut2 = cutoff * cutoff;
    uint32_t at = offsets[i];
    for (int ky = b.ky0; ky <= b.ky1; ++ky) {
        const uint32_t s0 = sg.bucket_first[(size_t)ky * sg.buckets_x + b.kx0];
        const uint32_t s1 = sg.bucket_first[(size_t)ky * sg.buckets_x + b.kx1 + 1];
        for (uint32_t s = s0; s < s1; ++s) {
            const double dx = sg.sources[s].x_m - rx, dy = sg.sources[s].y_m - ry;
            if (dx * dx + dy * dy > cut2) continue;
            pair_receiver[at] = r;
            pair_source[at] = s;
            ++at;
        }
    }
}

__global__ __launch_bounds__(256) void features_kernel(Scene sc, const qm_receiver* receivers, const qm_source* sources,
                                                       const uint32_t* pair_receiver, const uint32_t* pair_source,
                                                       uint32_t n, uint16_t* rows, uint8_t* in_envelope,
                                                       float* source_alt, uint32_t* exact_list,
                                                       uint32_t* exact_count) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    Crossing cross[CROSS_CAP];
    const qm_receiver& rcv = receivers[pair_receiver[i]];
    const qm_source& src = sources[pair_source[i]];
    float f[FEATURES];
    const SurrogatePath sp = surrogate_features(sc, rcv, src, cross, f);
    uint16_t* x0 = rows + (size_t)i * 2 * FEATURES;
    for (int k = 0; k < FEATURES; ++k) {
        const uint16_t h = to_bf16(f[k]);
        x0[k] = h;
        x0[FEATURES + k] = k == 37 ? to_bf16(1.0f) : h;
    }
    in_envelope[i] = sp.in_envelope ? 1 : 0;
    if (sp.in_envelope) {
        // The source's altitude for the energy terms (its ground as the exact path samples it).
        const qm_terrain& T = sc.terrain;
        const double inv = 1.0 / T.spacing_m;
        GroundSample s0;
        ground_at(T, (src.x_m - T.origin_x_m) * inv, (src.y_m - T.origin_y_m) * inv, s0);
        source_alt[i] = s0.z + src.height_m;
    } else {
        exact_list[atomicAdd(exact_count, 1u)] = i;
    }
}

__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void exact_kernel(
    Scene sc, BandTable B, const qm_receiver* receivers, const qm_source* sources, const uint32_t* pair_receiver,
    const uint32_t* pair_source, const uint32_t* exact_list, const uint32_t* exact_count, float* exact_energy,
    uint8_t* exact_failed) {
    const uint32_t count = *exact_count;
    PathScratch w;
    StateStream states[2];
    w.state = states;
    for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < count; k += gridDim.x * blockDim.x) {
        const uint32_t i = exact_list[k];
        PathResult r;
        evaluate_path(sc, B, receivers[pair_receiver[i]], sources[pair_source[i]], w, r);
        for (int q = 0; q < PERIODS; ++q) {
            const float s01 = r.bands[q][0] + r.bands[q][1], s23 = r.bands[q][2] + r.bands[q][3];
            const float s45 = r.bands[q][4] + r.bands[q][5], s67 = r.bands[q][6] + r.bands[q][7];
            exact_energy[(size_t)i * PERIODS + q] = r.failed ? 0.0f : (s01 + s23) + (s45 + s67);
        }
        exact_failed[i] = r.failed ? 1 : 0;
    }
}

// popup/src/evaluate.rs:119-203 and physics/src/ray.rs:155-190 as the exact path applies them (no
// foliage: the envelope has none), with the network's per-band A_ground + A_dif of both states.
__global__ __launch_bounds__(256) void combine_kernel(const qm_receiver* receivers, uint32_t first, uint32_t n,
                                                      const qm_source* sources, const uint32_t* offsets,
                                                      const uint32_t* pair_source, const uint8_t* in_envelope,
                                                      const float* source_alt, const float* att,
                                                      const float* exact_energy, const uint8_t* exact_failed,
                                                      double* period_energy, uint32_t* paths, uint32_t* failures,
                                                      uint32_t* surrogate_paths) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t r = first + i;
    const qm_receiver& rcv = receivers[r];
    double energy[PERIODS] = {0.0, 0.0, 0.0};
    uint32_t fails = 0, learned = 0;
    const uint32_t p0 = offsets[i], p1 = offsets[i + 1];
    for (uint32_t p = p0; p < p1; ++p) {
        float e[PERIODS];
        if (!in_envelope[p]) {
            if (exact_failed[p]) {
                ++fails;
                continue;
            }
            for (int q = 0; q < PERIODS; ++q) e[q] = exact_energy[(size_t)p * PERIODS + q];
        } else {
            ++learned;
            const qm_source& src = sources[pair_source[p]];
            const double dxd = rcv.x_m - src.x_m, dyd = rcv.y_m - src.y_m;
            const double length_d = hypot(dxd, dyd);
            const float H = (float)fmax(length_d, 1.0);
            const float dz = rcv.altitude_m - source_alt[p];
            const float slant = fmaxf(sqrtf(H * H + dz * dz), 1.0f);
            const float azimuth = atan2f((float)dyd, (float)dxd);
            float pf[PERIODS];
            for (int q = 0; q < PERIODS; ++q) pf[q] = favourable_at(rcv.favourable[q], azimuth);
            const float hd = fmaxf((float)length_d, src.exclusion_radius_m);
            const float d = fmaxf(sqrtf(hd * hd + dz * dz), 1.0f);
            const float weight = 1.0f / (d * d * 12.589254117941673f);
            const float* a0 = att + (size_t)p * 2 * BANDS;
            float bands[PERIODS][BANDS];
            for (int b = 0; b < BANDS; ++b) {
                const float air = __expf(-rcv.alpha_db_per_km[b] * slant * 1e-3f * (LN10 / 10.0f));
                const float eh = air * __expf(-a0[b] * (LN10 / 10.0f));
                const float ef = air * __expf(-a0[BANDS + b] * (LN10 / 10.0f));
                for (int q = 0; q < PERIODS; ++q)
                    bands[q][b] = weight * src.energy[q][b] * (pf[q] * ef + (1.0f - pf[q]) * eh);
            }
            for (int q = 0; q < PERIODS; ++q) {
                const float s01 = bands[q][0] + bands[q][1], s23 = bands[q][2] + bands[q][3];
                const float s45 = bands[q][4] + bands[q][5], s67 = bands[q][6] + bands[q][7];
                e[q] = (s01 + s23) + (s45 + s67);
            }
        }
        for (int q = 0; q < PERIODS; ++q) energy[q] += (double)e[q];
    }
    for (int q = 0; q < PERIODS; ++q) period_energy[(size_t)r * PERIODS + q] = energy[q];
    paths[r] = p1 - p0;
    failures[r] = fails;
    if (surrogate_paths) surrogate_paths[r] = learned;
}

__global__ void zero_u32(uint32_t* p) { *p = 0u; }

namespace {

int fail(char* err, size_t errlen, const char* fmt, ...) {
    if (err && errlen) {
        va_list args;
        va_start(args, fmt);
        vsnprintf(err, errlen, fmt, args);
        va_end(args);
    }
    return 1;
}

#define SP_TRY(expr)                                                                   \
    do {                                                                               \
        const hipError_t e_ = (expr);                                                  \
        if (e_ != hipSuccess) return fail(err, errlen, "%s: %s", #expr, hipGetErrorString(e_)); \
    } while (0)

unsigned blocks_for(size_t n) { return (unsigned)((n + 255) / 256); }

// Device allocations freed on the stream when the call returns (success or not).
struct StreamAllocations {
    hipStream_t stream;
    std::vector<void*> ptrs;
    ~StreamAllocations() {
        for (void* p : ptrs) (void)hipFreeAsync(p, stream);
    }
    hipError_t alloc(void** p, size_t bytes) {
        const hipError_t e = hipMallocAsync(p, std::max<size_t>(bytes, 256), stream);
        if (e == hipSuccess) ptrs.push_back(*p);
        return e;
    }
};

// Pairs a chunk of the surrogate painter holds at most.
constexpr size_t CHUNK_PAIRS = 8u << 20;

}  // namespace

extern "C" int qm_surrogate_samples_device(qm_gpu* gpu, const qm_square* square, const qm_receiver* d_receivers,
                                           const uint32_t* d_pair_receiver, const uint32_t* d_pair_source,
                                           size_t n_pairs, float* d_inputs, float* d_targets, uint32_t* d_in_envelope,
                                           void* stream, char* err, size_t errlen) {
    if (!gpu || !square || (n_pairs && (!d_receivers || !d_pair_receiver || !d_pair_source || !d_inputs ||
                                        !d_targets || !d_in_envelope)))
        return fail(err, errlen, "null argument");
    if (n_pairs == 0) return 0;
    if (n_pairs > 0xffffffffu) return fail(err, errlen, "too many pairs");
    SP_TRY(hipSetDevice(gpu->ordinal));
    hipLaunchKernelGGL(samples_kernel, dim3(blocks_for(n_pairs)), dim3(256), 0, (hipStream_t)stream, square->scene,
                       gpu->bands, d_receivers, square->d_sources_by_input, d_pair_receiver, d_pair_source,
                       (uint32_t)n_pairs, d_inputs, d_targets, d_in_envelope);
    SP_TRY(hipGetLastError());
    return 0;
}

extern "C" int qm_surrogate_paint_receivers_device(qm_gpu* gpu, const qm_square* square, qm_mlp* mlp,
                                                   const qm_receiver* d_receivers, uint32_t nx, uint32_t ny,
                                                   double cutoff_m, double* d_period_energy, uint32_t* d_paths,
                                                   uint32_t* d_failures, uint32_t* d_surrogate_paths, void* stream_ptr,
                                                   char* err, size_t errlen) {
    if (!gpu || !square || !mlp || !d_receivers || !d_period_energy || !d_paths || !d_failures)
        return fail(err, errlen, "null argument");
    if (!(cutoff_m >= 0.0)) return fail(err, errlen, "cutoff must be non-negative");
    const size_t n = (size_t)nx * ny;
    if (n == 0) return 0;
    if (n > 0xffffffffu) return fail(err, errlen, "too many receivers");
    SP_TRY(hipSetDevice(gpu->ordinal));
    const hipStream_t s = (hipStream_t)stream_ptr;
    StreamAllocations mem{s, {}};
    // Pairs per receiver, then receiver chunks of at most CHUNK_PAIRS pairs (on the host).
    uint32_t* d_counts = nullptr;
    SP_TRY(mem.alloc((void**)&d_counts, n * 4));
    hipLaunchKernelGGL(count_kernel, dim3(blocks_for(n)), dim3(256), 0, s, d_receivers, (uint32_t)n, square->grid,
                       cutoff_m, d_counts);
    SP_TRY(hipGetLastError());
    std::vector<uint32_t> counts(n);
    SP_TRY(hipMemcpyAsync(counts.data(), d_counts, n * 4, hipMemcpyDeviceToHost, s));
    SP_TRY(hipStreamSynchronize(s));
    size_t largest = 0;
    for (size_t i = 0; i < n; ++i) largest = std::max<size_t>(largest, counts[i]);
    if (largest > CHUNK_PAIRS) return fail(err, errlen, "a receiver has %zu sources in reach (at most %zu)", largest, CHUNK_PAIRS);
    uint32_t *d_offsets, *d_pair_receiver, *d_pair_source, *d_exact_list, *d_exact_count;
    uint16_t* d_rows;
    uint8_t *d_in_envelope, *d_exact_failed;
    float *d_source_alt, *d_att, *d_exact_energy;
    const size_t P = CHUNK_PAIRS;
    SP_TRY(mem.alloc((void**)&d_offsets, (n + 1) * 4));
    SP_TRY(mem.alloc((void**)&d_pair_receiver, P * 4));
    SP_TRY(mem.alloc((void**)&d_pair_source, P * 4));
    SP_TRY(mem.alloc((void**)&d_exact_list, P * 4));
    SP_TRY(mem.alloc((void**)&d_exact_count, 4));
    SP_TRY(mem.alloc((void**)&d_rows, P * 2 * FEATURES * 2));
    SP_TRY(mem.alloc((void**)&d_in_envelope, P));
    SP_TRY(mem.alloc((void**)&d_exact_failed, P));
    SP_TRY(mem.alloc((void**)&d_source_alt, P * 4));
    SP_TRY(mem.alloc((void**)&d_att, P * 2 * BANDS * 4));
    SP_TRY(mem.alloc((void**)&d_exact_energy, P * PERIODS * 4));
    std::vector<uint32_t> offsets;
    int cus = 0;
    SP_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, gpu->ordinal));
    size_t first = 0;
    while (first < n) {
        // The chunk: receivers first .. last-1 with at most CHUNK_PAIRS pairs.
        offsets.assign(1, 0u);
        size_t last = first, pairs = 0;
        while (last < n && pairs + counts[last] <= P) {
            pairs += counts[last];
            offsets.push_back((uint32_t)pairs);
            ++last;
        }
        const uint32_t m = (uint32_t)(last - first);
        SP_TRY(hipMemcpyAsync(d_offsets, offsets.data(), offsets.size() * 4, hipMemcpyHostToDevice, s));
        hipLaunchKernelGGL(list_kernel, dim3(blocks_for(m)), dim3(256), 0, s, d_receivers, (uint32_t)first, m,
                           square->grid, cutoff_m, d_offsets, d_pair_receiver, d_pair_source);
        hipLaunchKernelGGL(zero_u32, dim3(1), dim3(1), 0, s, d_exact_count);
        SP_TRY(hipGetLastError());
        if (pairs) {
            hipLaunchKernelGGL(features_kernel, dim3(blocks_for(pairs)), dim3(256), 0, s, square->scene, d_receivers,
                               square->grid.sources, d_pair_receiver, d_pair_source, (uint32_t)pairs, d_rows,
                               d_in_envelope, d_source_alt, d_exact_list, d_exact_count);
            SP_TRY(hipGetLastError());
            if (qm::mlp_infer_rows(mlp, d_rows, pairs * 2, d_att, s, err, errlen)) return 1;
            hipLaunchKernelGGL(exact_kernel, dim3((unsigned)std::max(cus, 1) * 4), dim3(256), 0, s, square->scene,
                               gpu->bands, d_receivers, square->grid.sources, d_pair_receiver, d_pair_source,
                               d_exact_list, d_exact_count, d_exact_energy, d_exact_failed);
            SP_TRY(hipGetLastError());
        }
        hipLaunchKernelGGL(combine_kernel, dim3(blocks_for(m)), dim3(256), 0, s, d_receivers, (uint32_t)first, m,
                           square->grid.sources, d_offsets, d_pair_source, d_in_envelope, d_source_alt, d_att,
                           d_exact_energy, d_exact_failed, d_period_energy, d_paths, d_failures, d_surrogate_paths);
        SP_TRY(hipGetLastError());
        first = last;
    }
    return 0;
}
