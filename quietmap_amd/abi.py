"""ctypes mirrors of the scene ABI in ``csrc/qm_scene.h`` and numpy record dtypes of its arrays.

The scene is handed over as plain pointers and sizes, as ``paint-gpu/src/square_upload.rs:18-287``
hands the reference's card its flat buffers; these classes hold numpy arrays alive and expose the
C structs that point into them.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

BANDS = 8
PERIODS = 3
SECTORS = 16
LAYERS = 6
PAIR_SUMS = 9
CELL_STEPS = 256
RECEIVER_HEIGHT_M = 4.0
GROUND_FROM_TERRAIN = 255
WEATHER_NODE_BYTES = 64
KIND_BUILDING = 0
KIND_BARRIER = 1

FAILED_NOT_READ = 1
FAILED_NO_TERRAIN = 2
FAILED_CAPACITY = 4

RUN_DTYPE = np.dtype([("outline", "<u4"), ("first_edge", "<u2"), ("edge_count", "<u2")])
OUTLINE_DTYPE = np.dtype(
    [
        ("footprint_id", "<u8"),
        ("first_vertex", "<u4"),
        ("vertex_count", "<u4"),
        ("height_dm", "<u2"),
        ("kind", "u1"),
        ("pad", "u1", (5,)),
    ]
)
SOURCE_DTYPE = np.dtype(
    [
        ("x_m", "<f8"),
        ("y_m", "<f8"),
        ("height_m", "<f4"),
        ("platform_half_width_m", "<f4"),
        ("exclusion_radius_m", "<f4"),
        ("ground_percent", "u1"),
        ("layer", "u1"),
        ("on_deck", "u1"),
        ("pad", "u1"),
        ("footprint_id", "<u8"),
        ("energy", "<f4", (PERIODS, BANDS)),
        ("deck_m", "<f4"),
        ("pad2", "<u4"),
    ]
)
POINT_DTYPE = np.dtype([("x_m", "<f8"), ("y_m", "<f8"), ("own_footprint", "<u8")])
# qm_line_piece (qm_scene.h): a straight line-source piece and its emission per metre.
LINE_PIECE_DTYPE = np.dtype([("ax_m", "<f8"), ("ay_m", "<f8"), ("bx_m", "<f8"), ("by_m", "<f8"),
                             ("height_m", "<f4"), ("layer", "u1"), ("on_deck", "u1"), ("pad", "u1", 2),
                             ("energy_per_m", "<f4", (3, 8)), ("deck_a_m", "<f4"), ("deck_b_m", "<f4")])
assert LINE_PIECE_DTYPE.itemsize == 144
# qm_receiver (qm_scene.h): a point's receiver properties as the card computes them.
RECEIVER_DTYPE = np.dtype([("x_m", "<f8"), ("y_m", "<f8"), ("own_footprint", "<u8"), ("altitude_m", "<f4"),
                           ("failed", "<u4"), ("alpha_db_per_km", "<f4", (8,)), ("favourable", "<f4", (3, 16))])
assert RECEIVER_DTYPE.itemsize == 256
# qm_tile (qm_tiles.h): a tile's receivers (row-major nx x ny from receiver_first) and its sources.
TILE_DTYPE = np.dtype([("receiver_first", "<u4"), ("nx", "<u4"), ("ny", "<u4"), ("source_first", "<u4"),
                       ("source_count", "<u4")])
assert RUN_DTYPE.itemsize == 8 and OUTLINE_DTYPE.itemsize == 24
assert SOURCE_DTYPE.itemsize == 144 and POINT_DTYPE.itemsize == 24


def _ptr(array, ctype):
    if array is None:
        return C.cast(None, C.POINTER(ctype))
    return array.ctypes.data_as(C.POINTER(ctype))


class CTerrain(C.Structure):
    _fields_ = [
        ("origin_x_m", C.c_double),
        ("origin_y_m", C.c_double),
        ("spacing_m", C.c_double),
        ("columns", C.c_uint32),
        ("rows", C.c_uint32),
        ("nodes", C.c_void_p),
        ("canopy_m", C.c_void_p),
        ("fine_detail", C.c_void_p),
        ("fine_blocks", C.c_void_p),
        ("column_m", C.c_double),
        ("row_y_m", C.c_void_p),
    ]


FINE_PER_NODE = 6
DETAIL_STEP_M = 0.2
FINE_ENDS_M = 100.0
FINE_BLOCK_CELLS = 32


class CObstacles(C.Structure):
    _fields_ = [
        ("origin_steps_x", C.c_double),
        ("origin_steps_y", C.c_double),
        ("metres_per_step", C.c_double),
        ("cell_x0", C.c_int32),
        ("cell_y0", C.c_int32),
        ("cells_x", C.c_uint32),
        ("cells_y", C.c_uint32),
        ("runs_count", C.c_uint32),
        ("outlines_count", C.c_uint32),
        ("vertices_count", C.c_uint32),
        ("pad", C.c_uint32),
        ("cell_first_run", C.c_void_p),
        ("runs", C.c_void_p),
        ("outlines", C.c_void_p),
        ("vertices", C.c_void_p),
        ("metres_per_step_y", C.c_double),
    ]


class CWeather(C.Structure):
    _fields_ = [
        ("origin_x_m", C.c_double),
        ("origin_y_m", C.c_double),
        ("spacing_m", C.c_double),
        ("columns", C.c_uint32),
        ("rows", C.c_uint32),
        ("nodes", C.c_void_p),
        ("nodes_per_degree", C.c_uint32),
        ("pad", C.c_uint32),
        ("frame_x", C.c_double),
        ("frame_y", C.c_double),
        ("east_m_per_unit", C.c_double),
        ("north_m_per_unit", C.c_double),
    ]


class Terrain:
    """A terrain lattice: ``nodes`` (rows, columns) u32 = height code | imperviousness << 16 |
    forest << 24, ``canopy_m`` (rows, columns) u8; row r at y = origin_y + r*spacing. Optional
    fine terrain: ``fine_detail`` (FINE_PER_NODE*(rows-1)+1, FINE_PER_NODE*(columns-1)+1) i16 steps
    of DETAIL_STEP_M and ``fine_blocks`` (block rows, block columns) u8 flags of FINE_BLOCK_CELLS
    square blocks. The geographic frame: ``row_y_m`` (rows ascending north metres of each row) and
    ``column_m`` (east metres between columns), both from the square's Mercator frame."""

    def __init__(self, origin_x_m, origin_y_m, spacing_m, nodes, canopy_m=None, fine_detail=None, fine_blocks=None,
                 column_m=0.0, row_y_m=None):
        self.nodes = np.ascontiguousarray(nodes, dtype=np.uint32)
        rows, columns = self.nodes.shape
        self.canopy_m = (
            np.zeros((rows, columns), np.uint8) if canopy_m is None else np.ascontiguousarray(canopy_m, np.uint8)
        )
        self.origin_x_m, self.origin_y_m, self.spacing_m = float(origin_x_m), float(origin_y_m), float(spacing_m)
        self.fine_detail = self.fine_blocks = None
        if fine_detail is not None:
            self.fine_detail = np.ascontiguousarray(fine_detail, np.int16)
            assert self.fine_detail.shape == (FINE_PER_NODE * (rows - 1) + 1, FINE_PER_NODE * (columns - 1) + 1)
            self.fine_blocks = np.ascontiguousarray(fine_blocks, np.uint8)
            assert self.fine_blocks.shape == ((rows - 2) // FINE_BLOCK_CELLS + 1, (columns - 2) // FINE_BLOCK_CELLS + 1)
        self.column_m = float(column_m)
        self.row_y_m = None
        if row_y_m is not None:
            self.row_y_m = np.ascontiguousarray(row_y_m, np.float64)
            assert self.row_y_m.shape == (rows,) and np.all(np.diff(self.row_y_m) > 0)
        self.c = CTerrain(
            self.origin_x_m,
            self.origin_y_m,
            self.spacing_m,
            columns,
            rows,
            self.nodes.ctypes.data,
            self.canopy_m.ctypes.data,
            None if self.fine_detail is None else self.fine_detail.ctypes.data,
            None if self.fine_blocks is None else self.fine_blocks.ctypes.data,
            self.column_m,
            None if self.row_y_m is None else self.row_y_m.ctypes.data,
        )

    @staticmethod
    def pack(height_m, impervious_percent, forest_percent=None):
        """Nodes from heights (m, coded at 0.2 m as tiles/src/terrain.rs:41-45) and percentages."""
        code = np.clip(np.round((np.asarray(height_m, np.float64) + 500.0) * 5.0), 0, 65534).astype(np.uint32)
        imp = np.asarray(impervious_percent).astype(np.uint32)
        forest = np.zeros_like(imp) if forest_percent is None else np.asarray(forest_percent).astype(np.uint32)
        return code | (imp << 16) | (forest << 24)


class Obstacles:
    """The obstacle cell index (see qm_scene.h)."""

    def __init__(self, origin_steps, metres_per_step, cell_origin, cells, cell_first_run, runs, outlines, vertices,
                 metres_per_step_y=0.0):
        self.cell_first_run = np.ascontiguousarray(cell_first_run, np.uint32)
        self.runs = np.ascontiguousarray(runs, RUN_DTYPE)
        self.outlines = np.ascontiguousarray(outlines, OUTLINE_DTYPE)
        self.vertices = np.ascontiguousarray(vertices, np.int32).reshape(-1, 2)
        self.origin_steps = (float(origin_steps[0]), float(origin_steps[1]))
        self.metres_per_step = float(metres_per_step)
        self.cell_origin = (int(cell_origin[0]), int(cell_origin[1]))
        self.cells = (int(cells[0]), int(cells[1]))
        assert self.cell_first_run.size == self.cells[0] * self.cells[1] + 1
        self.c = CObstacles(
            self.origin_steps[0],
            self.origin_steps[1],
            self.metres_per_step,
            self.cell_origin[0],
            self.cell_origin[1],
            self.cells[0],
            self.cells[1],
            self.runs.size,
            self.outlines.size,
            self.vertices.shape[0],
            0,
            self.cell_first_run.ctypes.data,
            self.runs.ctypes.data if self.runs.size else None,
            self.outlines.ctypes.data if self.outlines.size else None,
            self.vertices.ctypes.data if self.vertices.size else None,
            float(metres_per_step_y),
        )
        self.metres_per_step_y = float(metres_per_step_y)


class Weather:
    """A weather lattice of 64-byte nodes (physics/src/weather.rs:19-27)."""

    def __init__(self, origin_x_m, origin_y_m, spacing_m, nodes, frame=None, nodes_per_degree=0):
        """``frame`` (Mercator x, y of the frame's origin in z12 units, east and north metres per
        unit) with ``nodes_per_degree``: the world's table read by latitude and longitude."""
        self.nodes = np.ascontiguousarray(nodes, np.uint8)
        rows, columns, width = self.nodes.shape
        assert width == WEATHER_NODE_BYTES
        self.frame = frame
        self.nodes_per_degree = int(nodes_per_degree)
        fx, fy, em, nm = frame if frame is not None else (0.0, 0.0, 0.0, 0.0)
        if self.nodes_per_degree:
            assert rows == 180 * self.nodes_per_degree + 1 and columns == 360 * self.nodes_per_degree
        self.c = CWeather(float(origin_x_m), float(origin_y_m), float(spacing_m), columns, rows, self.nodes.ctypes.data,
                          self.nodes_per_degree, 0, float(fx), float(fy), float(em), float(nm))

    @staticmethod
    def node(percent, alpha_db_per_km):
        """One node's bytes: percent [3][16] (0-100) and alpha per band (dB/km, coded in 0.01)."""
        out = np.zeros(WEATHER_NODE_BYTES, np.uint8)
        out[: PERIODS * SECTORS] = np.asarray(percent, np.uint8).reshape(-1)
        codes = np.round(np.asarray(alpha_db_per_km, np.float64) * 100.0).astype("<u2")
        out[PERIODS * SECTORS :] = codes.view(np.uint8)
        return out


class Scene:
    """Terrain, obstacles and weather of one painted area."""

    def __init__(self, terrain: Terrain, obstacles: Obstacles, weather: Weather):
        self.terrain, self.obstacles, self.weather = terrain, obstacles, weather

# Aircraft (csrc/qm_aircraft.h).
NPD_DISTANCES = 10
POWER_ROWS = 4
INSTALL_WING, INSTALL_FUSELAGE, INSTALL_PROPELLER = 0, 1, 2
HIST_BINS = 160
HIST_FLOOR_DB = 40.0
HIST_STEP_DB = 0.5
PASS_CLASSES = 48
PASS_CLASS_DB = 3.0
PASS_LOWEST_DB = -30.0
PASS_FLOATS = 2 * 3 * PASS_CLASSES
HORIZON_BANDS = 6
TERRAIN_SECTORS = 32
BUILDING_SECTORS = 256
TERRAIN_HORIZON_FLOATS = 1 + TERRAIN_SECTORS * (1 + 2 * HORIZON_BANDS)
HORIZON_FLOATS = TERRAIN_HORIZON_FLOATS + 1 + BUILDING_SECTORS * (1 + 2 * HORIZON_BANDS)
NO_HORIZON = 0xFFFFFFFF


class AircraftScreening(C.Structure):
    """qm_aircraft_screening (csrc/qm_aircraft.h)."""
    _fields_ = [("horizons", C.c_void_p), ("n_horizons", C.c_size_t), ("horizon_of", C.c_void_p),
                ("vertex_ground_m", C.c_void_p)]
AIRCRAFT_CLASS_DTYPE = np.dtype(
    [
        ("sel_db", "<f8", (2, POWER_ROWS, NPD_DISTANCES)),
        ("lamax_db", "<f8", (2, POWER_ROWS, NPD_DISTANCES)),
        ("rows", "<u4", (2,)),
        ("installation", "<u4"),
        ("pad", "<u4"),
        ("v_ref_kt", "<f8"),
    ]
)
TRACK_DTYPE = np.dtype(
    [
        ("first_vertex", "<u4"),
        ("vertex_count", "<u4"),
        ("aircraft_class", "<u4"),
        ("departure", "<u4"),
        ("power_row", "<u4"),
        ("power_weight", "<f4"),
        ("speed_kt", "<f4"),
        ("flights", "<u4", (3,)),
    ]
)
assert AIRCRAFT_CLASS_DTYPE.itemsize == 2 * 2 * POWER_ROWS * NPD_DISTANCES * 8 + 24
assert TRACK_DTYPE.itemsize == 40

# Aircraft boxes at the painter's points (qm_aircraft.h qm_flight_box / qm_flight_point /
# qm_flight_segment; the reference card's FlightBox / FlightPoint / FlightSegment).
FLIGHT_SUMS = 3 + 2 * 3 * PASS_CLASSES
FLIGHT_WEIGHT_ONE, FLIGHT_WEIGHT_FAR, FLIGHT_WEIGHT_NEAR = 0, 1, 2
AIRCRAFT_REACH_M = 16000.0
NO_HORIZONS = 0xFFFFFFFFFFFFFFFF
FLIGHT_BOX_DTYPE = np.dtype(
    [
        ("east_m", "<f8"),
        ("north_m", "<f8"),
        ("centroid_altitude_m", "<f8"),
        ("sin_axis", "<f8"),
        ("cos_axis", "<f8"),
        ("gradient", "<f8"),
        ("gradient_spread", "<f8"),
        ("average_length_m", "<f8"),
        ("passing", "<f8", (3,)),
        ("piece_length_m", "<f4", (3,)),
        ("levels_db", "<f4", (3, NPD_DISTANCES)),
        ("tail_levels_db", "<f4", (3,)),
        ("lg_scaled_distance", "<f4", (NPD_DISTANCES,)),
        ("installation_shares", "<f4", (3, 2, 3)),
        ("screened", "<u4"),
        ("helicopter", "<u4"),
        ("silent", "<u4", (3,)),
        ("pad", "<u4"),
    ]
)
FLIGHT_POINT_DTYPE = np.dtype(
    [("x", "<f8"), ("y", "<f8"), ("altitude_m", "<f8"), ("horizons", "<u8"), ("impedance_gain", "<f4"),
     ("pad", "<u4")]
)
FLIGHT_SEGMENT_DTYPE = np.dtype(
    [("within_m", "<f8"), ("point", "<u4"), ("sum", "<u4"), ("first", "<u4"), ("count", "<u4"), ("weight", "<u4"),
     ("coarse_m", "<f4")]
)
assert FLIGHT_BOX_DTYPE.itemsize == 368
assert FLIGHT_POINT_DTYPE.itemsize == 40
assert FLIGHT_SEGMENT_DTYPE.itemsize == 32
