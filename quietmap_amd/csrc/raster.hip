// libqmpaint: raster-out on the card (include/qm_paint.h qm_level_cells, qm_hm3_predict): a painted
// square's quantized cells and the HM3 tiles' predicted streams, the byte-exact steps before the
// host's Brotli.
//
//   cells_kernel  - per pixel (paint/src/levels.rs:58-66 PaintedCells::from_painted): Nden cells of
//                   each layer's and the place's loudness (levels.rs:26-34, twice the loudness level
//                   in phon), Lden cells of each layer's energy and of their sum (levels.rs:87-109,
//                   twice Lden in dB), 254 for silence, 255 for a pixel not assessed; an invalid
//                   energy is flagged for the host to refuse, as lden_cell's Err.
//   paeth_kernel  - per cell of a 512 x 512 tile (paint/src/hm3.rs:68-108 encode_metric_tile before
//                   Brotli): the metric's 5-byte header and the layer byte, then each cell minus
//                   PNG's Paeth predictor of it (filter 4, its left / above / corner tie order),
//                   read from the tile's cells; a lane a cell, the rows coalesced.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>

#include "../../include/qm_paint.h"
#include "qm_internal.hpp"

namespace {

constexpr int LAYERS = 6;
constexpr int CELLS_A_PIXEL = LAYERS + 1;
constexpr uint8_t NO_LEVEL = 255;
constexpr uint8_t SILENCE = 254;
constexpr int TILE = QM_HM3_TILE_PX;

int rfail(char* err, size_t errlen, const char* fmt, ...) {
    if (err && errlen) {
        va_list args;
        va_start(args, fmt);
        vsnprintf(err, errlen, fmt, args);
        va_end(args);
    }
    return 1;
}

// levels.rs cell: 80 + 20 log2(sone) rounded half away from zero, SILENCE when NaN or below 0.
__device__ __forceinline__ uint8_t nden_cell(double sone) {
    const double byte = 80.0 + 20.0 * log2(sone);
    if (isnan(byte) || byte < 0.0) return SILENCE;
    return (uint8_t)fmin(round(byte), (double)(SILENCE - 1));
}

// levels.rs lden_cell: 2 x 10 lg(energy) rounded, SILENCE under 1 (0 dB); `bad` for an energy the
// reference refuses (negative or not finite).
__device__ __forceinline__ uint8_t lden_cell(double energy, bool& bad) {
    if (!isfinite(energy) || energy < 0.0) {
        bad = true;
        return SILENCE;
    }
    if (energy < 1.0) return SILENCE;
    return (uint8_t)fmin(round(2.0 * (10.0 * log10(energy))), (double)(SILENCE - 1));
}

__global__ __launch_bounds__(256) void cells_kernel(const uint8_t* __restrict__ assessed,
                                                    const double* __restrict__ layer_energy,
                                                    const double* __restrict__ ndens, size_t pixels,
                                                    uint8_t* __restrict__ nden_out, uint8_t* __restrict__ lden_out,
                                                    unsigned long long* first_bad) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= pixels) return;
    uint8_t nd[CELLS_A_PIXEL], ld[CELLS_A_PIXEL];
    if (!assessed[i]) {
        for (int k = 0; k < CELLS_A_PIXEL; ++k) nd[k] = ld[k] = NO_LEVEL;
    } else {
        bool bad = false;
        double total = 0.0;
        for (int k = 0; k < LAYERS; ++k) {
            const double e = layer_energy[i * LAYERS + k];
            bool layer_bad = false;
            ld[k] = lden_cell(e, layer_bad);
            if (layer_bad && !bad) {
                bad = true;
                atomicMin(first_bad, (unsigned long long)i * CELLS_A_PIXEL + k);
            }
            total += e;
        }
        bool total_bad = false;
        ld[LAYERS] = lden_cell(total, total_bad);
        if (total_bad && !bad) atomicMin(first_bad, (unsigned long long)i * CELLS_A_PIXEL + LAYERS);
        for (int k = 0; k < CELLS_A_PIXEL; ++k) nd[k] = nden_cell(ndens[i * CELLS_A_PIXEL + k]);
    }
    for (int k = 0; k < CELLS_A_PIXEL; ++k) {
        nden_out[i * CELLS_A_PIXEL + k] = nd[k];
        lden_out[i * CELLS_A_PIXEL + k] = ld[k];
    }
}

// hm3.rs paeth_prediction: the tie order left, above, corner.
__device__ __forceinline__ uint8_t paeth(int left, int above, int corner) {
    const int estimate = left + above - corner;
    const int dl = abs(estimate - left), da = abs(estimate - above), dc = abs(estimate - corner);
    if (dl <= da && dl <= dc) return (uint8_t)left;
    if (da <= dc) return (uint8_t)above;
    return (uint8_t)corner;
}

// A tile a blockIdx.y, a row of 512 cells a blockIdx.x (two cells a work-item).
__global__ __launch_bounds__(256) void paeth_kernel(const uint8_t* __restrict__ cells,
                                                    const uint8_t* __restrict__ tile_layer, int metric,
                                                    uint8_t* __restrict__ out) {
    const size_t tile = blockIdx.y;
    const int row = blockIdx.x;
    const uint8_t* c = cells + tile * (size_t)TILE * TILE;
    uint8_t* o = out + tile * (size_t)QM_HM3_PREDICTED_BYTES;
    if (row == 0 && threadIdx.x == 0) {
        const char* header = metric == QM_HM3_LDEN ? "HMD \x01" : "HM3 \x07";
        for (int k = 0; k < 5; ++k) o[k] = (uint8_t)header[k];
        o[5] = tile_layer[tile];
    }
    for (int x = threadIdx.x; x < TILE; x += blockDim.x) {
        const size_t at = (size_t)row * TILE + x;
        const int left = x > 0 ? c[at - 1] : 0;
        const int above = row > 0 ? c[at - TILE] : 0;
        const int corner = x > 0 && row > 0 ? c[at - TILE - 1] : 0;
        o[6 + at] = (uint8_t)(c[at] - paeth(left, above, corner));
    }
}

struct Device {
    void* p = nullptr;
    ~Device() {
        if (p) (void)hipFree(p);
    }
};

}  // namespace

extern "C" int qm_level_cells_device(qm_gpu* gpu, const uint8_t* d_assessed, const double* d_layer_energy,
                                     const double* d_ndens, size_t pixels, uint8_t* d_nden_cells,
                                     uint8_t* d_lden_cells, unsigned long long* d_first_bad, void* stream, char* err,
                                     size_t errlen) {
    if (!gpu || (pixels && (!d_assessed || !d_layer_energy || !d_ndens || !d_nden_cells || !d_lden_cells ||
                            !d_first_bad)))
        return rfail(err, errlen, "null argument");
    if (pixels == 0) return 0;
    hipError_t e = hipSetDevice(gpu->ordinal);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(cells_kernel, dim3((unsigned)((pixels + 255) / 256)), dim3(256), 0, (hipStream_t)stream,
                           d_assessed, d_layer_energy, d_ndens, pixels, d_nden_cells, d_lden_cells, d_first_bad);
        e = hipGetLastError();
    }
    if (e != hipSuccess) return rfail(err, errlen, "level cells: %s", hipGetErrorString(e));
    return 0;
}

extern "C" int qm_level_cells(qm_gpu* gpu, const uint8_t* assessed, const double* layer_energy, const double* ndens,
                              size_t pixels, uint8_t* nden_cells, uint8_t* lden_cells, char* err, size_t errlen) {
    if (!gpu || (pixels && (!assessed || !layer_energy || !ndens || !nden_cells || !lden_cells)))
        return rfail(err, errlen, "null argument");
    if (pixels == 0) return 0;
    if (pixels > ((size_t)1 << 40)) return rfail(err, errlen, "%zu pixels exceed one call", pixels);
    Device d_assessed, d_energy, d_ndens, d_nden, d_lden, d_bad;
    hipError_t e = hipSetDevice(gpu->ordinal);
    if (e == hipSuccess) e = hipMalloc(&d_assessed.p, pixels);
    if (e == hipSuccess) e = hipMalloc(&d_energy.p, sizeof(double) * pixels * LAYERS);
    if (e == hipSuccess) e = hipMalloc(&d_ndens.p, sizeof(double) * pixels * CELLS_A_PIXEL);
    if (e == hipSuccess) e = hipMalloc(&d_nden.p, pixels * CELLS_A_PIXEL);
    if (e == hipSuccess) e = hipMalloc(&d_lden.p, pixels * CELLS_A_PIXEL);
    if (e == hipSuccess) e = hipMalloc(&d_bad.p, sizeof(unsigned long long));
    const unsigned long long none = ~0ull;
    if (e == hipSuccess) e = hipMemcpy(d_assessed.p, assessed, pixels, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(d_energy.p, layer_energy, sizeof(double) * pixels * LAYERS, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(d_ndens.p, ndens, sizeof(double) * pixels * CELLS_A_PIXEL, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(d_bad.p, &none, sizeof(none), hipMemcpyHostToDevice);
    if (e != hipSuccess) return rfail(err, errlen, "level cells: %s", hipGetErrorString(e));
    if (qm_level_cells_device(gpu, (const uint8_t*)d_assessed.p, (const double*)d_energy.p, (const double*)d_ndens.p,
                              pixels, (uint8_t*)d_nden.p, (uint8_t*)d_lden.p, (unsigned long long*)d_bad.p, nullptr,
                              err, errlen))
        return 1;
    unsigned long long bad = none;
    e = hipMemcpy(&bad, d_bad.p, sizeof(bad), hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(nden_cells, d_nden.p, pixels * CELLS_A_PIXEL, hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(lden_cells, d_lden.p, pixels * CELLS_A_PIXEL, hipMemcpyDeviceToHost);
    if (e != hipSuccess) return rfail(err, errlen, "level cells: %s", hipGetErrorString(e));
    if (bad != none) {
        const size_t pixel = (size_t)(bad / CELLS_A_PIXEL), slot = (size_t)(bad % CELLS_A_PIXEL);
        if (slot < (size_t)LAYERS)
            return rfail(err, errlen, "pixel %zu, layer %zu: invalid Lden energy %g", pixel, slot,
                         layer_energy[pixel * LAYERS + slot]);
        return rfail(err, errlen, "pixel %zu, total: invalid Lden energy", pixel);
    }
    return 0;
}

extern "C" int qm_hm3_predict_device(qm_gpu* gpu, const uint8_t* d_cells, const uint8_t* d_tile_layer, size_t tiles,
                                     int metric, uint8_t* d_out, void* stream, char* err, size_t errlen) {
    if (!gpu || (tiles && (!d_cells || !d_tile_layer || !d_out))) return rfail(err, errlen, "null argument");
    if (metric != QM_HM3_NDEN && metric != QM_HM3_LDEN) return rfail(err, errlen, "unknown metric %d", metric);
    if (tiles == 0) return 0;
    if (tiles > 65535) return rfail(err, errlen, "%zu tiles exceed one launch", tiles);
    hipError_t e = hipSetDevice(gpu->ordinal);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(paeth_kernel, dim3(TILE, (unsigned)tiles), dim3(256), 0, (hipStream_t)stream, d_cells,
                           d_tile_layer, metric, d_out);
        e = hipGetLastError();
    }
    if (e != hipSuccess) return rfail(err, errlen, "hm3 predict: %s", hipGetErrorString(e));
    return 0;
}

extern "C" int qm_hm3_predict(qm_gpu* gpu, const uint8_t* cells, const uint8_t* tile_layer, size_t tiles, int metric,
                              uint8_t* out, char* err, size_t errlen) {
    if (!gpu || (tiles && (!cells || !tile_layer || !out))) return rfail(err, errlen, "null argument");
    for (size_t t = 0; t < tiles; ++t)
        if (tile_layer[t] > LAYERS) return rfail(err, errlen, "tile %zu: unknown HM3 layer %u", t, tile_layer[t]);
    if (tiles == 0) return 0;
    const size_t in_bytes = tiles * (size_t)TILE * TILE, out_bytes = tiles * (size_t)QM_HM3_PREDICTED_BYTES;
    Device d_cells, d_layer, d_out;
    hipError_t e = hipSetDevice(gpu->ordinal);
    if (e == hipSuccess) e = hipMalloc(&d_cells.p, in_bytes);
    if (e == hipSuccess) e = hipMalloc(&d_layer.p, tiles);
    if (e == hipSuccess) e = hipMalloc(&d_out.p, out_bytes);
    if (e == hipSuccess) e = hipMemcpy(d_cells.p, cells, in_bytes, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(d_layer.p, tile_layer, tiles, hipMemcpyHostToDevice);
    if (e != hipSuccess) return rfail(err, errlen, "hm3 predict: %s", hipGetErrorString(e));
    if (qm_hm3_predict_device(gpu, (const uint8_t*)d_cells.p, (const uint8_t*)d_layer.p, tiles, metric,
                              (uint8_t*)d_out.p, nullptr, err, errlen))
        return 1;
    e = hipMemcpy(out, d_out.p, out_bytes, hipMemcpyDeviceToHost);
    if (e != hipSuccess) return rfail(err, errlen, "hm3 predict: %s", hipGetErrorString(e));
    return 0;
}
