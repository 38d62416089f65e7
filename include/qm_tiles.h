/* libqmpaint: the batched painter of many uneven tiles (one persistent launch in a hipGraph).
 *
 * The reference paints one square at a time, 2M pairs per launch (paint-gpu/src/batch.rs:22,
 * device.rs:440-765); here thousands of tiles, each a receiver grid painted against its own
 * sources only, share one persistent work-stealing launch: the work is cut into items (a tile's
 * 16x16 receiver block against a block of its sources), queued heaviest tile first and striped
 * over the XCDs, and the launch sequence (clearing the outputs and the queue heads with fill
 * kernels, then the painter) is captured once in a hipGraph that qm_tiles_run replays.
 * The source block size and the groups per CU are the plan's parameters; quietmap_amd/autotune.py
 * picks them by timing candidates on a sample of the tiles.
 *
 * Conventions as qm_paint.h.
 */
#ifndef QM_TILES_H
#define QM_TILES_H

#include <stddef.h>
#include <stdint.h>

#include "qm_paint.h"

#ifdef __cplusplus
extern "C" {
#endif

/* A tile: receivers receiver_first .. receiver_first + nx*ny - 1 (row-major, x fastest) of the
 * receiver array, painted against sources source_first .. source_first + source_count - 1 of the
 * square's sources in upload order. */
typedef struct qm_tile {
    uint32_t receiver_first;
    uint32_t nx, ny;
    uint32_t source_first, source_count;
} qm_tile;

typedef struct qm_tile_plan qm_tile_plan;

/* A plan for `tiles` over the square and the device receivers (`d_receivers`, n_receivers as
 * qm_receivers_device makes them): work items of `block_sources` sources (>= 1), a persistent
 * launch of `groups_per_cu` 256-work-item groups per CU, outputs per receiver into the caller's
 * device arrays (period energy, 3 doubles; paths; failed paths), all captured in a hipGraph. */
int qm_tiles_plan(qm_gpu* gpu, const qm_square* square, const qm_tile* tiles, size_t n_tiles,
                  const qm_receiver* d_receivers, size_t n_receivers, uint32_t block_sources, uint32_t groups_per_cu,
                  double* d_period_energy, uint32_t* d_paths, uint32_t* d_failures, qm_tile_plan** out, char* err,
                  size_t errlen);
/* Replays the plan's graph on `stream` (outputs cleared, then painted). */
int qm_tiles_run(qm_tile_plan* plan, void* stream, char* err, size_t errlen);
/* Paths (receiver-source pairs) one run paints, and the plan's work items. */
uint64_t qm_tiles_paths(const qm_tile_plan* plan);
uint64_t qm_tiles_items(const qm_tile_plan* plan);
void qm_tiles_plan_free(qm_tile_plan* plan);

#ifdef __cplusplus
}
#endif

#endif
