/* The scene as the painter reads it: plain C structs of pointers and sizes, shared by the host
 * library, the HIP kernels and (as a data layout only) the FP64 oracle.
 *
 * Frame: planar metres, x east, y north. Three parts, each mirroring what the reference uploads
 * for a square (paint-gpu/src/square_upload.rs:18-287):
 *
 *  - terrain: one regular lattice of decoded nodes, 4 bytes each as the reference's decoded
 *    qmterr2 node (tiles/src/terrain.rs:21-27, 149-159): height code u16 (-500 m + code/5),
 *    imperviousness u8 (G = 1 - imperviousness/100), forest cover u8, plus a canopy-height byte
 *    plane. Row r lies at y = origin_y + r*spacing, column c at x = origin_x + c*spacing; in the
 *    geographic frame (a square's one arc-second lattice, popup/src/scene.rs) row r lies at
 *    y = row_y_m[r] (uneven: latitude through Mercator) and column c at origin_x + c*column_m.
 *  - obstacles: the cell index of buildings and walls in integer "steps" (y south), cells of 256
 *    steps, per cell the runs of consecutive edges of one outline through it, outlines with their
 *    footprint id, height (dm) and kind, vertices as exact integers (tiles/src/obstacles.rs:1-23,
 *    popup/src/obstacles/mod.rs:1-8). Steps of a frame point: sx = origin_steps_x + x/mps,
 *    sy = origin_steps_y - y/mps.
 *  - weather: a lattice of the reference's 64-byte weather nodes (physics/src/weather.rs:19-27):
 *    48 favourable percentages [period][sector], then 8 u16 air absorptions in 0.01 dB/km,
 *    bilinear between nodes.
 */
#ifndef QM_SCENE_H
#define QM_SCENE_H

#include <stdint.h>

#define QM_BANDS 8
#define QM_PERIODS 3
#define QM_SECTORS 16
#define QM_LAYERS 6
#define QM_WEATHER_NODE_BYTES 64
#define QM_CELL_STEPS 256
#define QM_RECEIVER_HEIGHT_M 4.0
#define QM_GROUND_FROM_TERRAIN 255
#define QM_PAIR_SUMS 9

/* Per-pair failure bits (paint-gpu/kernels/common.cuh:27-31 semantics). */
#define QM_FAILED_NOT_READ 1u   /* the ray leaves the scene's obstacle cells */
#define QM_FAILED_NO_TERRAIN 2u /* a terrain node outside the lattice */
#define QM_FAILED_CAPACITY 4u   /* a per-path buffer (crossings, hull, roofs) too small */
/* With QM_FAILED_CAPACITY, which buffer (diagnostic detail; callers may mask it off). */
#define QM_DETAIL_CROSSINGS 16u
#define QM_DETAIL_ROOFS 32u
#define QM_DETAIL_HULL 64u
#define QM_DETAIL_PROFILE 128u
#define QM_DETAIL_WALK 256u

#define QM_KIND_BUILDING 0u
#define QM_KIND_BARRIER 1u

#ifdef __cplusplus
extern "C" {
#endif

/* Fine terrain (tiles/src/terrain_fine.rs): the lattice QM_FINE_PER_NODE times as fine, detail
 * over the coarse bilinear ground in steps of QM_DETAIL_STEP_M, read within QM_FINE_ENDS_M of a
 * ray's ends (physics/src/profile.rs:13) where a block of QM_FINE_BLOCK_CELLS square coarse cells
 * under that stretch has fine terrain (the reference's tiles with a fine_offset). */
#define QM_FINE_PER_NODE 6
#define QM_DETAIL_STEP_M 0.2
#define QM_FINE_ENDS_M 100.0
#define QM_FINE_BLOCK_CELLS 32

typedef struct qm_terrain {
    double origin_x_m, origin_y_m;
    double spacing_m;
    uint32_t columns, rows;
    const uint32_t* nodes;   /* rows*columns: height code | impervious << 16 | forest << 24 */
    const uint8_t* canopy_m; /* rows*columns */
    /* Optional fine terrain (null: none): detail steps on (QM_FINE_PER_NODE (rows-1) + 1) x
     * (QM_FINE_PER_NODE (columns-1) + 1) fine nodes, row-major from the south-west node, and a
     * flag per block ((columns-2)/QM_FINE_BLOCK_CELLS + 1 blocks a row, likewise for rows). */
    const int16_t* fine_detail;
    const uint8_t* fine_blocks;
    /* The geographic frame (0 / null: the square lattice above). `column_m`: east metres between
     * columns; `row_y_m`: rows ascending values, the frame north metres of each row (the rows of
     * the one arc-second lattice; DeviceSquare::upload's scene_rows, paint-gpu/src/
     * square_upload.rs:40-56). Fine rows lie evenly between the table's. */
    double column_m;
    const double* row_y_m;
} qm_terrain;

typedef struct qm_run {
    uint32_t outline;
    uint16_t first_edge;
    uint16_t edge_count;
} qm_run;

typedef struct qm_outline {
    uint64_t footprint_id;
    uint32_t first_vertex;  /* into the vertex array */
    uint32_t vertex_count;  /* a ring repeats its first vertex last */
    uint16_t height_dm;
    uint8_t kind;           /* QM_KIND_* */
    uint8_t pad[5];
} qm_outline;

typedef struct qm_obstacles {
    double origin_steps_x, origin_steps_y;
    double metres_per_step;
    int32_t cell_x0, cell_y0; /* global cell of grid cell (0, 0) */
    uint32_t cells_x, cells_y;
    uint32_t runs_count, outlines_count, vertices_count, pad;
    const uint32_t* cell_first_run; /* cells_x*cells_y + 1, row-major by cell y */
    const qm_run* runs;
    const qm_outline* outlines;
    const int32_t* vertices; /* 2 per vertex: x, y steps */
    /* North metres per step where they differ from the east's (the geographic frame's
     * LocalFrame, tiles/src/geo.rs:224-243: east and north metres per Mercator unit on the
     * WGS84 ellipsoid); 0: metres_per_step both ways. */
    double metres_per_step_y;
} qm_obstacles;

typedef struct qm_weather {
    double origin_x_m, origin_y_m;
    double spacing_m;
    uint32_t columns, rows;
    const uint8_t* nodes; /* rows*columns*64 */
    /* The geographic frame (nodes_per_degree 0: the planar lattice above). The nodes are then the
     * world's table (physics/src/weather.rs:15-25: rows 180 npd + 1 from 90 N, columns 360 npd
     * from 0 E, wrapping) and a place is read by its latitude and longitude (weather.rs:127-135,
     * 186-200), the frame mapping metres to z12 Mercator units (geo::Mercator: 4096 a side;
     * x = frame_x + east/east_m_per_unit, y = frame_y - north/north_m_per_unit; paint-gpu/kernels/
     * paint.cu:35-41). */
    uint32_t nodes_per_degree, pad;
    double frame_x, frame_y, east_m_per_unit, north_m_per_unit;
} qm_weather;

/* A point source: the candidate's place and its attribute (popup/src/candidates.rs:35-49). A
 * source `on_deck` stands on a bridge's deck at `deck_m` (the candidate's ground0, the same datum
 * as the terrain) instead of the terrain under it (physics/src/ray.rs RayEnds::source_ground_m:
 * its altitude is the deck's plus its height, and the terrain within its platform is cut to the
 * deck). */
typedef struct qm_source {
    double x_m, y_m;
    float height_m;
    float platform_half_width_m;
    float exclusion_radius_m;
    uint8_t ground_percent; /* QM_GROUND_FROM_TERRAIN: the terrain's G under the source */
    uint8_t layer;
    uint8_t on_deck;
    uint8_t pad;
    uint64_t footprint_id;
    float energy[QM_PERIODS][QM_BANDS]; /* A-weighted sound energy per period and band */
    float deck_m;
    uint32_t pad2;
} qm_source;

/* A straight line-source piece: a line candidate's ends and attribute (popup/src/candidates.rs:
 * 35-49 with `line` set; physics/src/line.rs integrates it). Ends in frame metres, the source's
 * height above the ground under it, the emission per metre of the piece. A piece `on_deck` runs
 * on a bridge's deck from `deck_a_m` under its first end to `deck_b_m` under its second (the
 * candidate's ground0 and ground1), each node standing on the deck between them. */
typedef struct qm_line_piece {
    double ax_m, ay_m, bx_m, by_m;
    float height_m;
    uint8_t layer;
    uint8_t on_deck;
    uint8_t pad[2];
    float energy_per_m[QM_PERIODS][QM_BANDS]; /* A-weighted sound energy per metre */
    float deck_a_m, deck_b_m;
} qm_line_piece;

/* A receiver point (paint/src/exact.rs:18-34): where it stands and the building it stands in. */
typedef struct qm_point {
    double x_m, y_m;
    uint64_t own_footprint;
} qm_point;

/* Receiver properties as the device computes them (paint-gpu kernels point_receivers). */
typedef struct qm_receiver {
    double x_m, y_m;
    uint64_t own_footprint;
    float altitude_m;
    uint32_t failed;
    float alpha_db_per_km[QM_BANDS];
    float favourable[QM_PERIODS][QM_SECTORS];
} qm_receiver;

#ifdef __cplusplus
}
#endif

#endif
