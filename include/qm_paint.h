/* libqmpaint: the MI355X noise painter's C ABI.
 *
 * The drop-in boundary for the reference's paint-gpu crate (paint-gpu/src/lib.rs:14-22): the
 * Rust side keeps reading tiles and calls these functions through `extern "C"`; the library owns
 * everything on the device. Every call returns 0 on success, or a nonzero status with a message
 * in `err` (at most `errlen` bytes, NUL-terminated) that the Rust wrapper maps to Err(String).
 * Per-pair soft failures are not errors: they come back as QM_FAILED_* bits per pair (as
 * paint-gpu/kernels/common.cuh:27-31), for the caller's CPU fallback (paint-gpu/src/batch.rs:292-317).
 *
 * Ownership: input arrays are borrowed for the duration of the call; handles are freed by their
 * *_free function. Threading: a qm_gpu may be used from several threads; each call uses its own
 * stream and buffers. Scene layout: qm_scene.h (in quietmap_amd/csrc/).
 */
#ifndef QM_PAINT_H
#define QM_PAINT_H

#include <stddef.h>
#include <stdint.h>

#include "../quietmap_amd/csrc/qm_aircraft.h"
#include "../quietmap_amd/csrc/qm_scene.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct qm_gpu qm_gpu;
typedef struct qm_square qm_square;
typedef struct qm_flight_boxes qm_flight_boxes;

/* Gpu::new(ordinal) (paint-gpu/src/device.rs:254-420): opens the device and checks that it is a
 * gfx950 with the library's kernels loaded. */
int qm_gpu_new(int ordinal, qm_gpu** out, char* err, size_t errlen);
/* Gpu::name (paint-gpu/src/device.rs:422-425). */
int qm_gpu_name(const qm_gpu* gpu, char* out, size_t len);
void qm_gpu_free(qm_gpu* gpu);

/* DeviceSquare::upload (paint-gpu/src/square_upload.rs:18-287): the scene's terrain lattice,
 * obstacle cell index, weather lattice and point sources (the candidates with their attributes),
 * copied to the device. The sources are bucketed on a grid of `bucket_m` metres for the painter. */
int qm_square_upload(qm_gpu* gpu, const qm_terrain* terrain, const qm_obstacles* obstacles,
                     const qm_weather* weather, const qm_source* sources, size_t n_sources, double bucket_m,
                     qm_square** out, char* err, size_t errlen);
/* DeviceSquare::bytes (paint-gpu/src/square_upload.rs:283-287): device bytes the square holds. */
size_t qm_square_bytes(const qm_square* square);
void qm_square_free(qm_square* square);

/* Gpu::evaluate_pairs (paint-gpu/src/device.rs:440-765): for every pair (point, source) the
 * received energy per period mixed by the weather, then calm and downwind alone
 * (`sums`, 9 floats a pair: mixed[3], calm[3], favourable[3]; paint/src/exact.rs:129-133), its
 * failure bits, every point's bands per layer, period and band (`point_bands`, optional,
 * n_points*6*3*8 floats, summed with the optional per-pair `pair_weight`; exact.rs:208-216), and
 * each pair's own mixed bands (`pair_bands`, optional, n_pairs*3*8 floats, unweighted, zero for a
 * failed pair: the `flat` output hum_pairs.rs reads for Batch::pair_spectra). Host arrays. */
int qm_evaluate_pairs(qm_gpu* gpu, const qm_square* square, const qm_point* points, size_t n_points,
                      const uint32_t* pair_point, const uint32_t* pair_source, const float* pair_weight,
                      size_t n_pairs, float* sums, uint32_t* failed, float* point_bands, float* pair_bands, char* err,
                      size_t errlen);

/* The painted grid: receivers on a row-major grid (`points`, nx*ny, x fastest), every source
 * within `cutoff_m` (horizontal) of each, in one persistent work-stealing launch. Out per
 * receiver: the summed mixed energy per period (`period_energy`, 3 doubles; Lden follows from
 * physics/src/bands.rs:59-66), the paths evaluated and the paths that failed. Host arrays. */
int qm_paint_grid(qm_gpu* gpu, const qm_square* square, const qm_point* points, uint32_t nx, uint32_t ny,
                  double cutoff_m, double* period_energy, uint32_t* paths, uint32_t* failures, char* err,
                  size_t errlen);

/* The same with device arrays (already resident in HBM) on a caller's stream (hipStream_t, or
 * NULL for the default stream); returns after enqueueing. */
int qm_paint_grid_device(qm_gpu* gpu, const qm_square* square, const qm_point* d_points, uint32_t nx, uint32_t ny,
                         double cutoff_m, double* d_period_energy, uint32_t* d_paths, uint32_t* d_failures,
                         void* stream, char* err, size_t errlen);

/* No counterpart in the reference. qm_paint_grid_device for callers that cannot allocate device
 * memory: code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc. It
 * allocates nothing: the receiver records and queue heads live in a static device array of the
 * library (QM_NOALLOC_RECEIVERS records of 256 bytes, 16 MiB, plus 256 bytes), one per device,
 * held until the process ends. More than QM_NOALLOC_RECEIVERS receivers (nx*ny) fail the call.
 * Every call on a device shares that device's array, so calls that overlap on one device race:
 * order them on one stream (or with events). Enqueued on `stream`; returns after enqueueing. */
#define QM_NOALLOC_RECEIVERS 65536
int qm_paint_grid_device_noalloc(qm_gpu* gpu, const qm_square* square, const qm_point* d_points, uint32_t nx,
                                 uint32_t ny, double cutoff_m, double* d_period_energy, uint32_t* d_paths,
                                 uint32_t* d_failures, void* stream, char* err, size_t errlen);

/* The two halves of qm_paint_grid_device. Receiver properties of points on the device: ground
 * altitude + 4 m, air absorption and favourable percentages (the reference's point_receivers,
 * paint-gpu/kernels/paint.cu:62, over paint/src/exact.rs:95-118). */
int qm_receivers_device(qm_gpu* gpu, const qm_square* square, const qm_point* d_points, size_t n,
                        qm_receiver* d_receivers, void* stream, char* err, size_t errlen);
/* The grid painter over prepared receivers (nx*ny, row-major). */
int qm_paint_receivers_device(qm_gpu* gpu, const qm_square* square, const qm_receiver* d_receivers, uint32_t nx,
                              uint32_t ny, double cutoff_m, double* d_period_energy, uint32_t* d_paths,
                              uint32_t* d_failures, void* stream, char* err, size_t errlen);

/* ISO 532-1:2017 stationary (Zwicker) loudness in sone of each cell's 28 third-octave levels
 * (25 Hz - 12.5 kHz, dB): the reference card's loud_zwicker (paint-gpu/kernels/loudness.cuh:85-153)
 * over physics/src/loudness.rs:148-182, the step inside Gpu::pixel_loudness
 * (paint-gpu/src/loudness.rs:270). Host arrays, `levels` cell-major (28 floats a cell). */
int qm_loudness(qm_gpu* gpu, const float* levels, size_t n, float* sone, char* err, size_t errlen);
/* The same on device arrays: level (cell i, band b) at d_levels[i*cell_stride + b*band_stride];
 * band-major planes (cell_stride 1, band_stride n) read coalesced. Enqueued on `stream`. */
int qm_loudness_device(qm_gpu* gpu, const float* d_levels, size_t n, size_t cell_stride, size_t band_stride,
                       float* d_sone, void* stream, char* err, size_t errlen);

/* Doc 29 aircraft noise at a receiver grid from raw tracks (quietmap_amd/csrc/qm_aircraft.h): the
 * reference card's Gpu::flight_sums (paint-gpu/src/aircraft.rs:235) evaluates the same segment
 * physics (physics/src/doc29/segment.rs:284-381) over aircraft boxes; this entry point takes the
 * flights' segments directly. Per receiver (`points_xy`, nx*ny x/y pairs, `receiver_altitude_m`
 * above the tracks' datum, every segment within `reach_m` horizontally): the SEL energy per period
 * summed over flights a day (`period_energy`, 3 doubles), the flights a day whose LAmax reaches
 * 65 and 70 dB, `hist` (QM_HIST_BINS counts a receiver: flights per 0.5 dB of LAmax from 40 dB)
 * and, unless null, `passes` (QM_PASS_FLOATS doubles a receiver): the pass classes of
 * FlightPasses::add_pass (popup/src/aircraft.rs:86-95, the card's pass_class and flight_sums
 * class sums, paint-gpu/kernels/aircraft.cuh:250-311), each heard segment a pass at its SEL energy
 * over the period's seconds while its flights are on it (box_passing, boxes.rs:225-265, with the
 * segment's 3-D length as the chord). With `screening` (null: unscreened), each segment's SEL
 * below the ceiling is screened by the receiver's horizons (segment.rs:284-350 over
 * screening.rs and horizons.rs; the reference card's horizon_loss_db, aircraft.cuh:83-100): the
 * terrain's loss at the line's closest point, the buildings' at the segment's, their larger beyond
 * the lateral attenuation; LAmax is not screened. Host arrays. */
int qm_aircraft_grid(qm_gpu* gpu, const qm_aircraft_class* classes, size_t n_classes, const qm_track* tracks,
                     size_t n_tracks, const float* vertices, size_t n_vertices, const double* points_xy, uint32_t nx,
                     uint32_t ny, double receiver_altitude_m, double reach_m, double* period_energy, uint32_t* n65,
                     uint32_t* n70, uint32_t* hist, double* passes, const qm_aircraft_screening* screening,
                     char* err, size_t errlen);

/* Pixel loudness on the card (the reference's Gpu::pixel_loudness, paint-gpu/src/loudness.rs, over
 * its kernels pixel_mixtures and pixel_ndens, paint-gpu/kernels/distribution.cuh and
 * loudness.cuh): per packed pixel (paint/src/sone.rs pack: per layer its steady energy per
 * period, its line count and each line's 43 floats; then per curve, the six layers' and the
 * place's, and period 28 third-octave levels at 0 dB(A), the first NaN for none; pixel i from
 * `offsets[i]` floats into `input`) the level distributions of every period, hour group and
 * weather state at the map's 2 dB resolution, mixed per period (`mixtures`, optional:
 * pixels x 3 periods x QM_LOUD_CURVES x QM_LOUD_BINS x 2 floats, a bin's share of the time and
 * share x mean intensity), and Nden per curve (`ndens`, QM_LOUD_CURVES floats a pixel: each layer
 * alone, then the place). `kurze` is the 41 x 100 quantile table (physics/src/percentile.rs
 * TABLE); null takes qm_kurze_table's. Malformed counts, and hourly profiles other than none
 * (-1) and 0-2, fail the call. Host arrays. */
#define QM_LOUD_BINS 90
#define QM_LOUD_CURVES 7
int qm_pixel_loudness(qm_gpu* gpu, const float* input, size_t input_floats, const uint64_t* offsets, size_t pixels,
                      const float* kurze, float* mixtures, float* ndens, char* err, size_t errlen);

/* The map's folding and packing of pixel sound on the card, then qm_pixel_loudness of the result
 * (the reference folds on the cores before its upload: paint/src/sone.rs:268-276 PixelSound::pack
 * over paint/src/sone_fold.rs:115-405 fold_placed_for_map, whose end is
 * popup/src/percentiles/weather.rs:139-175 merge_close_states at the map's spread 10^0.15, then
 * write_packed, sone.rs:217-234). `unfolded` holds pixels in qm_pixel_loudness's layout with every
 * value FP64 and the lines as placed: per layer, the weak events fold into one steady line within
 * their variance budget, the Kurze lines that never exceed 5 % of the held level into a steady line
 * per hourly profile, the passes that hold their mean likewise; the weather states each line keeps
 * within the spread merge; a layer's first steady line whose states read alike in FP32 joins its
 * steady energy. The folded pixel (FP32, at the pixel's offset, never longer than the unfolded one)
 * goes to `folded` when given (input_floats floats; NaN past each pixel's end), and on the card
 * into the loudness kernel; `mixtures` and `ndens` as qm_pixel_loudness's. Host arrays. */
int qm_pixel_loudness_unfolded(qm_gpu* gpu, const double* unfolded, size_t input_floats, const uint64_t* offsets,
                               size_t pixels, const float* kurze, float* mixtures, float* ndens, float* folded,
                               char* err, size_t errlen);

/* The quantile table of Kurze's line statistics (physics/src/percentile_table.rs): the
 * reference's seeded simulation (physics/src/percentile_tests.rs writes_the_table) restated, 41
 * rows of lambda from 1e-4 at 8 a decade by 100 quantiles, FP32. */
void qm_kurze_table(float* out);

/* Horizons::build and write_f32 of each point (popup/src/aircraft/horizons.rs:138-201; the
 * reference builds them on the host per card point, paint-gpu/src/aircraft.rs:252-258): the
 * terrain's and the buildings' QM_HORIZON_FLOATS floats a point, the screening input of
 * qm_aircraft_grid. At `altitude_m[i]` (null: the painter's receiver altitude, ground + 4 m),
 * the walls of the point's own footprint no horizon. `failed` per point: the receiver's failure
 * bits, a crossing search's (capacity) and QM_FAILED_NO_TERRAIN where a wall stands off the
 * terrain. Host arrays. */
int qm_horizons(qm_gpu* gpu, const qm_square* square, const qm_point* points, const double* altitude_m,
                size_t n_points, float* out, uint32_t* failed, char* err, size_t errlen);

/* DeviceFlightBoxes::upload (paint-gpu/src/aircraft.rs:160-205): a square's aircraft boxes as
 * the host writes them (qm_flight_box), copied to the device; freed by qm_flight_boxes_free. */
int qm_flight_boxes_upload(qm_gpu* gpu, const qm_flight_box* boxes, size_t n, qm_flight_boxes** out, char* err,
                           size_t errlen);
size_t qm_flight_boxes_count(const qm_flight_boxes* boxes);
void qm_flight_boxes_free(qm_flight_boxes* boxes);

/* Gpu::flight_sums (paint-gpu/src/aircraft.rs:222-350, kernels/aircraft.cuh:260-311): per segment
 * its boxes within reach, weighted, their SEL energy per period and their passes in the 48 pass
 * classes (popup/src/aircraft/boxes.rs:265-300 tile_flight_passes), added into sum slot
 * `segment.sum` of `sums` (n_sums x QM_FLIGHT_SUMS doubles, zeroed first). `horizons` holds the
 * points' horizon records (Horizons::write_f32 floats, QM_HORIZON_FLOATS from each point's
 * offset; QM_NO_HORIZONS: unscreened). Records out of range fail the call. Host arrays. */
int qm_flight_sums(qm_gpu* gpu, const qm_flight_boxes* boxes, const qm_flight_point* points, size_t n_points,
                   const float* horizons, size_t n_horizon_floats, const qm_flight_segment* segments,
                   size_t n_segments, double* sums, size_t n_sums, char* err, size_t errlen);
/* The same with device arrays on `stream` (a hipStream_t; null: the default stream), the sums
 * not zeroed (they accumulate); the caller keeps every index in range. */
int qm_flight_sums_device(qm_gpu* gpu, const qm_flight_boxes* boxes, const qm_flight_point* d_points,
                          const float* d_horizons, const qm_flight_segment* d_segments, size_t n_segments,
                          double* d_sums, void* stream, char* err, size_t errlen);

/* Gpu::evaluate_pairs for line candidates (paint-gpu/kernels/paint.cu:77-290 pair_rays, the
 * count / scan / list of rays, and ray_transfers; popup/src/evaluate.rs:171-267 source_rays): per
 * pair (point, line piece) the piece's quadrature nodes seen from the point (physics/src/line.rs,
 * the skyline mask of wide buckets), one exact ray per node, and the pair's 9 sums, failure bits,
 * rays (optional, per pair), point bands and pair bands (optional, as qm_evaluate_pairs). Host
 * arrays. */
int qm_evaluate_line_pairs(qm_gpu* gpu, const qm_square* square, const qm_point* points, size_t n_points,
                           const qm_line_piece* pieces, size_t n_pieces, const uint32_t* pair_point,
                           const uint32_t* pair_piece, size_t n_pairs, float* sums, uint32_t* failed, uint32_t* rays,
                           float* point_bands, float* pair_bands, char* err, size_t errlen);

/* The square's line candidates (DeviceSquare::upload keeps point and line candidates alike,
 * paint-gpu/src/square_upload.rs:172-220; popup/src/candidates.rs:89-104 Candidate with `line`):
 * line piece k is candidate n_sources + k of qm_evaluate_candidates. Replaces the lines set before;
 * ends must be finite and layers below QM_LAYERS. */
int qm_square_set_lines(qm_square* square, const qm_line_piece* pieces, size_t n_pieces, char* err, size_t errlen);
/* The square's candidates: its point sources, then its line pieces. */
size_t qm_square_candidates(const qm_square* square);

/* Gpu::evaluate_pairs as GpuBatch calls it (paint-gpu/src/device.rs:434-765, its argument
 * checks at 443-471): per pair its candidate (a point source or a line piece of the square) and
 * its receiver among `receivers`; per receiver its point in `bands` (`point_of_receiver`, null:
 * receiver r is point r). Out: each pair's 9 sums and failure bits (qm_evaluate_pairs' layout),
 * each pair's own mixed bands (`pair_spectra`, optional, n_pairs*3*8 floats, unweighted), and
 * `bands` (optional, n_points*6*3*8 floats) ADDED to: every successful pair's bands times its
 * `band_weight` (null: 1) at its receiver's point, layer by the candidate's (exact.rs:208-216).
 * A candidate outside the square, a receiver out of range or a point outside the bands fails the
 * call before any work. Host arrays. */
int qm_evaluate_candidates(qm_gpu* gpu, const qm_square* square, const qm_point* receivers, size_t n_receivers,
                           const uint32_t* point_of_receiver, size_t n_points, const uint32_t* candidate,
                           const uint32_t* receiver_of_pair, const float* band_weight, size_t n_pairs, float* sums,
                           uint32_t* failed, float* bands, float* pair_spectra, char* err, size_t errlen);

/* The crossing list of each segment (from, to) in frame metres, as the exact painter finds it
 * (popup/src/obstacles/crossings.rs:163-230; parity hook): per segment up to `capacity` crossings
 * (t, footprint, outline, edge index, flags bit 0 building / bit 1 bounds roof) and the count. */
int qm_crossings(qm_gpu* gpu, const qm_square* square, const double* segments, size_t n_segments,
                 uint32_t capacity, double* t, uint64_t* footprint, uint32_t* outline, uint32_t* edge,
                 uint32_t* flags, uint32_t* count, uint32_t* failed, char* err, size_t errlen);

/* Raster-out: a painted square's quantized cells (paint/src/levels.rs:58-109 PaintedCells::
 * from_painted): per pixel (`assessed` 0: 255 in every cell) the Nden cell of each layer's and
 * the place's loudness (`ndens`, 7 sone a pixel; levels.rs cell: 80 + 20 log2 sone, half a phon a
 * step, 254 below 0 phon) and the Lden cell of each layer's energy and of their sum (`layer_energy`,
 * 6 linear Lden energies a pixel; levels.rs lden_cell: twice Lden in dB, 254 below 0 dB), 7 bytes a
 * pixel each. A negative or non-finite energy fails the call naming the pixel and layer, as
 * lden_cell's Err does. Host arrays. */
int qm_level_cells(qm_gpu* gpu, const uint8_t* assessed, const double* layer_energy, const double* ndens,
                   size_t pixels, uint8_t* nden_cells, uint8_t* lden_cells, char* err, size_t errlen);
/* The same on device arrays, enqueued on `stream`: the lowest pixel x 7 + cell of an invalid energy
 * is atomically min-ed into `d_first_bad` (the caller sets it to ~0 first). */
int qm_level_cells_device(qm_gpu* gpu, const uint8_t* d_assessed, const double* d_layer_energy, const double* d_ndens,
                          size_t pixels, uint8_t* d_nden_cells, uint8_t* d_lden_cells, unsigned long long* d_first_bad,
                          void* stream, char* err, size_t errlen);

/* HM3 tiles before Brotli (paint/src/hm3.rs:68-108 encode_metric_tile; Brotli stays on the host):
 * per 512 x 512 tile of cells (row-major, tile after tile) the metric's 5-byte header ("HM3 \x07"
 * for Nden, "HMD \x01" for Lden), the tile's layer byte (`tile_layer`, 0-6) and each cell minus
 * PNG's Paeth predictor of it (filter 4 with its tie order), QM_HM3_PREDICTED_BYTES a tile. Host
 * arrays; a layer above 6 fails the call. */
#define QM_HM3_TILE_PX 512
#define QM_HM3_PREDICTED_BYTES (6 + QM_HM3_TILE_PX * QM_HM3_TILE_PX)
#define QM_HM3_NDEN 0
#define QM_HM3_LDEN 1
int qm_hm3_predict(qm_gpu* gpu, const uint8_t* cells, const uint8_t* tile_layer, size_t tiles, int metric,
                   uint8_t* out, char* err, size_t errlen);
/* The same on device arrays, enqueued on `stream` (layers unchecked: the caller's). */
int qm_hm3_predict_device(qm_gpu* gpu, const uint8_t* d_cells, const uint8_t* d_tile_layer, size_t tiles, int metric,
                          uint8_t* d_out, void* stream, char* err, size_t errlen);

#ifdef __cplusplus
}
#endif

#endif
