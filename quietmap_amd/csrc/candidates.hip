// Gpu::evaluate_pairs as the reference's batch calls it (paint-gpu/src/device.rs:434-765): pairs
// of any candidate kind in one call, point sources and line pieces of one square, each pair's
// receiver among the call's receivers and each receiver's point in the caller's band sums
// (`point_of_receiver`), bands added with the pairs' weights. The point pairs and the line pairs
// go to their kernels (painter.hip pairs_kernel, lines.hip's count / scan / list and
// line_rays_kernel) in two launches; the host maps receivers to points and puts every output back
// in the pairs' order.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <vector>

#include "../../include/qm_paint.h"
#include "qm_internal.hpp"

namespace {

int cfail(char* err, size_t errlen, const char* fmt, ...) {
    if (err && errlen) {
        va_list args;
        va_start(args, fmt);
        vsnprintf(err, errlen, fmt, args);
        va_end(args);
    }
    return 1;
}

constexpr size_t POINT_FLOATS = (size_t)QM_LAYERS * QM_PERIODS * QM_BANDS;
constexpr size_t PAIR_FLOATS = (size_t)QM_PERIODS * QM_BANDS;

}  // namespace

extern "C" int qm_square_set_lines(qm_square* square, const qm_line_piece* pieces, size_t n_pieces, char* err,
                                   size_t errlen) {
    if (!square || (n_pieces && !pieces)) return cfail(err, errlen, "null argument");
    if (square->n_sources + n_pieces > 0xffffffffu) return cfail(err, errlen, "too many candidates");
    for (size_t i = 0; i < n_pieces; ++i) {
        if (!std::isfinite(pieces[i].ax_m) || !std::isfinite(pieces[i].ay_m) || !std::isfinite(pieces[i].bx_m) ||
            !std::isfinite(pieces[i].by_m))
            return cfail(err, errlen, "piece %zu: ends not finite", i);
        if (pieces[i].layer >= QM_LAYERS) return cfail(err, errlen, "piece %zu: layer out of range", i);
    }
    square->lines.assign(pieces, pieces + n_pieces);
    return 0;
}

extern "C" size_t qm_square_candidates(const qm_square* square) {
    return square ? square->n_sources + square->lines.size() : 0;
}

extern "C" int qm_evaluate_candidates(qm_gpu* gpu, const qm_square* square, const qm_point* receivers,
                                      size_t n_receivers, const uint32_t* point_of_receiver, size_t n_points,
                                      const uint32_t* candidate, const uint32_t* receiver_of_pair,
                                      const float* band_weight, size_t n_pairs, float* sums, uint32_t* failed,
                                      float* bands, float* pair_spectra, char* err, size_t errlen) {
    if (!gpu || !square || (n_pairs && (!candidate || !receiver_of_pair || !sums || !failed)) ||
        (n_receivers && !receivers))
        return cfail(err, errlen, "null argument");
    const size_t n_candidates = square->n_sources + square->lines.size();
    for (size_t r = 0; r < n_receivers; ++r) {
        const size_t point = point_of_receiver ? point_of_receiver[r] : r;
        if (bands && point >= n_points)
            return cfail(err, errlen, "receiver %zu: point %zu outside the %zu points' bands", r, point, n_points);
    }
    // The two kinds' pairs, each kind's in the pairs' order.
    std::vector<uint32_t> point_pairs, line_pairs;
    for (size_t i = 0; i < n_pairs; ++i) {
        if (candidate[i] >= n_candidates)
            return cfail(err, errlen, "pair %zu: candidate %u outside the square's %zu", i, candidate[i], n_candidates);
        if (receiver_of_pair[i] >= n_receivers)
            return cfail(err, errlen, "pair %zu: receiver %u out of range", i, receiver_of_pair[i]);
        (candidate[i] < square->n_sources ? point_pairs : line_pairs).push_back((uint32_t)i);
    }
    std::vector<float> receiver_bands(bands ? n_receivers * POINT_FLOATS : 0, 0.0f);
    // Point sources: the weights go to the kernel, which sums each receiver's bands.
    if (!point_pairs.empty()) {
        const size_t n = point_pairs.size();
        std::vector<uint32_t> pp(n), ps(n);
        std::vector<float> pw(band_weight ? n : 0);
        for (size_t k = 0; k < n; ++k) {
            const uint32_t i = point_pairs[k];
            pp[k] = receiver_of_pair[i];
            ps[k] = candidate[i];
            if (band_weight) pw[k] = band_weight[i];
        }
        std::vector<float> s(n * QM_PAIR_SUMS), spectra(pair_spectra ? n * PAIR_FLOATS : 0);
        std::vector<uint32_t> f(n);
        if (qm_evaluate_pairs(gpu, square, receivers, n_receivers, pp.data(), ps.data(),
                              band_weight ? pw.data() : nullptr, n, s.data(), f.data(),
                              bands ? receiver_bands.data() : nullptr, pair_spectra ? spectra.data() : nullptr, err,
                              errlen))
            return 1;
        for (size_t k = 0; k < n; ++k) {
            const uint32_t i = point_pairs[k];
            std::memcpy(sums + (size_t)i * QM_PAIR_SUMS, &s[k * QM_PAIR_SUMS], sizeof(float) * QM_PAIR_SUMS);
            failed[i] = f[k];
            if (pair_spectra)
                std::memcpy(pair_spectra + (size_t)i * PAIR_FLOATS, &spectra[k * PAIR_FLOATS],
                            sizeof(float) * PAIR_FLOATS);
        }
    }
    // Line pieces: each pair's own bands come back and join its receiver's with its weight.
    if (!line_pairs.empty()) {
        const size_t n = line_pairs.size();
        std::vector<uint32_t> pp(n), pl(n);
        for (size_t k = 0; k < n; ++k) {
            const uint32_t i = line_pairs[k];
            pp[k] = receiver_of_pair[i];
            pl[k] = (uint32_t)(candidate[i] - square->n_sources);
        }
        std::vector<float> s(n * QM_PAIR_SUMS), spectra(n * PAIR_FLOATS);
        std::vector<uint32_t> f(n);
        if (qm_evaluate_line_pairs(gpu, square, receivers, n_receivers, square->lines.data(), square->lines.size(),
                                   pp.data(), pl.data(), n, s.data(), f.data(), nullptr, nullptr, spectra.data(), err,
                                   errlen))
            return 1;
        for (size_t k = 0; k < n; ++k) {
            const uint32_t i = line_pairs[k];
            std::memcpy(sums + (size_t)i * QM_PAIR_SUMS, &s[k * QM_PAIR_SUMS], sizeof(float) * QM_PAIR_SUMS);
            failed[i] = f[k];
            if (pair_spectra)
                std::memcpy(pair_spectra + (size_t)i * PAIR_FLOATS, &spectra[k * PAIR_FLOATS],
                            sizeof(float) * PAIR_FLOATS);
            if (bands && f[k] == 0) {
                const float w = band_weight ? band_weight[i] : 1.0f;
                const size_t layer = square->lines[pl[k]].layer;
                float* dst = &receiver_bands[((size_t)pp[k] * QM_LAYERS + layer) * PAIR_FLOATS];
                for (size_t b = 0; b < PAIR_FLOATS; ++b) dst[b] += w * spectra[k * PAIR_FLOATS + b];
            }
        }
    }
    // Each receiver's bands into its point's (added to what the caller's sums hold).
    if (bands)
        for (size_t r = 0; r < n_receivers; ++r) {
            const size_t point = point_of_receiver ? point_of_receiver[r] : r;
            float* dst = bands + point * POINT_FLOATS;
            const float* src = &receiver_bands[r * POINT_FLOATS];
            for (size_t b = 0; b < POINT_FLOATS; ++b) dst[b] += src[b];
        }
    return 0;
}
