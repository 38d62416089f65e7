/* Aircraft noise (ECAC Doc 29) as the painter reads it: noise classes with their NPD power rows,
 * flight tracks as polylines of segments, and the per-receiver outputs. Plain C structs shared by
 * the host library, the HIP kernels and (as a data layout only) the FP64 oracle.
 *
 * A class mirrors the reference's NPD profile and thrust rows (physics/src/doc29/npd.rs:62-96,
 * 256-330): SEL and LAmax at the ten NPD distances for up to QM_POWER_ROWS power settings per
 * operation (approach, departure), the NPD reference speed and the engine installation. A track is
 * one flight path flown `flights[period]` times a day: one class, one operation, one power
 * bracket (row and weight, physics/src/doc29/thrust.rs:104-119) and one ground speed.
 */
#ifndef QM_AIRCRAFT_H
#define QM_AIRCRAFT_H

#include <stddef.h>
#include <stdint.h>

#define QM_NPD_DISTANCES 10
#define QM_POWER_ROWS 4
#define QM_INSTALL_WING 0u
#define QM_INSTALL_FUSELAGE 1u
#define QM_INSTALL_PROPELLER 2u
/* The event-level histogram: LAmax bins of 0.5 dB from 40 dB; the last bin holds every level
 * above; levels below 40 dB are not binned. */
#define QM_HIST_BINS 160
#define QM_HIST_FLOOR_DB 40.0
#define QM_HIST_STEP_DB 0.5

/* Pass classes (popup/src/aircraft.rs:47-53): per period, 48 classes of 3 dB from -30 dB of the
 * level heard while a flight is on a segment, each the classes' summed energy (Leq, linear) and
 * their summed share of the period in progress: QM_PASS_FLOATS doubles a receiver, the energies
 * [period][class] then the passing shares [period][class]. */
#define QM_PASS_CLASSES 48
#define QM_PASS_CLASS_DB 3.0
#define QM_PASS_LOWEST_DB (-30.0)
#define QM_PASS_FLOATS (2 * 3 * QM_PASS_CLASSES)
#ifdef __cplusplus
static const double QM_PERIOD_HOURS[3] = {12.0, 4.0, 8.0};
#endif

/* Horizon screening (physics/src/doc29/screening.rs, popup/src/aircraft/horizons.rs): a
 * receiver's horizons as the reference writes them for its card (Horizons::write_f32): the
 * terrain's, then the buildings'; each the steepest tangent of all, then per sector (anticlockwise
 * from east) its steepest tangent and per range band the steepest edge's range (m) and tangent,
 * NaN range for an empty band. */
#define QM_HORIZON_BANDS 6
#define QM_TERRAIN_SECTORS 32
#define QM_BUILDING_SECTORS 256
#define QM_TERRAIN_HORIZON_FLOATS (1 + QM_TERRAIN_SECTORS * (1 + 2 * QM_HORIZON_BANDS))
#define QM_HORIZON_FLOATS (QM_TERRAIN_HORIZON_FLOATS + 1 + QM_BUILDING_SECTORS * (1 + 2 * QM_HORIZON_BANDS))
#define QM_NO_HORIZON 0xffffffffu
/* 10 lg 3 removed so the loss is 0 at grazing; 20 / 0.685 m a metre of path difference; the cap;
 * the height above the ground from which a segment is not screened. */
#define QM_GRAZING_LOSS_DB 4.771212547196624
#define QM_PATH_DIFFERENCE_SLOPE_PER_M 29.2
#define QM_SCREENING_CAP_DB 18.0
#define QM_SCREENING_CEILING_M 7620.0

#ifdef __cplusplus
extern "C" {
#endif

typedef struct qm_aircraft_class {
    double sel_db[2][QM_POWER_ROWS][QM_NPD_DISTANCES];   /* [departure][row][distance] */
    double lamax_db[2][QM_POWER_ROWS][QM_NPD_DISTANCES];
    uint32_t rows[2];                                   /* power rows per operation (1..4) */
    uint32_t installation;                              /* QM_INSTALL_* */
    double v_ref_kt;
} qm_aircraft_class;

/* Horizons for the aircraft grid: `horizons` holds `n_horizons` records of QM_HORIZON_FLOATS;
 * receiver i reads record horizon_of[i] (QM_NO_HORIZON: unscreened). `vertex_ground_m`, when not
 * null, is the terrain height under each track vertex (same datum as the vertices), else 0: a
 * segment at least QM_SCREENING_CEILING_M above the ground under its ends (their mean) is not
 * screened. */
typedef struct qm_aircraft_screening {
    const float* horizons;
    size_t n_horizons;
    const uint32_t* horizon_of;
    const float* vertex_ground_m;
} qm_aircraft_screening;

typedef struct qm_track {
    uint32_t first_vertex; /* into the vertex array: x, y, z float triples (frame metres, z above datum) */
    uint32_t vertex_count; /* segments = vertex_count - 1 */
    uint32_t aircraft_class;
    uint32_t departure;    /* 1 departure NPDs, 0 approach */
    uint32_t power_row;    /* the bracket's lower row */
    float power_weight;    /* toward the next row, in [0, 1] */
    float speed_kt;        /* ground speed (Delta_V) */
    uint32_t flights[3];   /* flights a day: day, evening, night */
} qm_track;

/* Aircraft boxes at the painter's points: the records the reference's host writes for its card's
 * flight_sums (paint-gpu/kernels/aircraft.cuh:14-50, filled by paint-gpu/src/aircraft.rs
 * DeviceFlightBoxes::upload and Gpu::flight_sums), field for field. A box (a tile's
 * tiles::aircraft::AircraftBox in the square's frame): its centroid east and north (frame
 * metres) and altitude, its axis, gradient and the gradients' spread, the average piece's length
 * (physics/src/doc29/boxes.rs average_length_m), the share of each period its flights are in it
 * (popup/src/aircraft/boxes.rs box_passing), per period the pieces' mean horizontal length, the
 * summed SEL at the ten NPD distances (dB, -inf silent) and at the tail anchor, lg d_lambda at the
 * NPD distances, the installation shares [period][at 1,000 ft, at the tail anchor][wing, fuselage,
 * propeller], whether it is screened (centroid below QM_SCREENING_CEILING_M over its ground), a
 * helicopter box (no receiver impedance) and each period's silence. */
#define QM_FLIGHT_SUMS (3 + 2 * 3 * QM_PASS_CLASSES)
#define QM_FLIGHT_WEIGHT_ONE 0u
#define QM_FLIGHT_WEIGHT_FAR 1u
#define QM_FLIGHT_WEIGHT_NEAR 2u
#define QM_AIRCRAFT_REACH_M 16000.0
/* A point whose `horizons` is QM_NO_HORIZONS is heard unscreened. */
#define QM_NO_HORIZONS UINT64_MAX

typedef struct qm_flight_box {
    double east_m, north_m;
    double centroid_altitude_m;
    double sin_axis, cos_axis;
    double gradient, gradient_spread;
    double average_length_m;
    double passing[3];
    float piece_length_m[3];
    float levels_db[3][QM_NPD_DISTANCES];
    float tail_levels_db[3];
    float lg_scaled_distance[QM_NPD_DISTANCES];
    float installation_shares[3][2][3];
    uint32_t screened;
    uint32_t helicopter;
    uint32_t silent[3];
} qm_flight_box;

/* A point: where it stands (frame metres) and its altitude, the offset of its horizon record
 * (QM_HORIZON_FLOATS floats) in the horizon array, and the NPD levels' impedance gain at its
 * elevation (energy, physics/src/doc29/atmosphere.rs receiver_impedance_db). */
typedef struct qm_flight_point {
    double x, y;
    double altitude_m;
    uint64_t horizons;
    float impedance_gain;
} qm_flight_point;

/* One point's boxes of one tile, file and band: `count` boxes from `first`, heard within
 * `within_m` (and QM_AIRCRAFT_REACH_M), each weighted by `weight` (QM_FLIGHT_WEIGHT_*, the
 * smoothstep from 1.5 to 3 `coarse_m`) at its horizontal distance, added to sum slot `sum`
 * (QM_FLIGHT_SUMS doubles: energy per period, then per period the pass classes' energies, then
 * their passing shares). */
typedef struct qm_flight_segment {
    double within_m;
    uint32_t point;
    uint32_t sum;
    uint32_t first;
    uint32_t count;
    uint32_t weight;
    float coarse_m;
} qm_flight_segment;

#ifdef __cplusplus
}
#endif

#endif
