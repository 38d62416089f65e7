"""Seeded synthetic scenes of the painter's benchmark configurations, emitted directly in the scene
ABI of ``csrc/qm_scene.h``: terrain lattice, obstacle cell index, weather lattice, point sources and
receiver points.

- ``road_scene`` (configuration "plain CNOSSOS road scene"): 256x256 receivers at 10 m, 4 m above
  flat ground with G per 10 m node from {0, 0.5, 1}; 200 random road polylines split into 2,000
  point sources at 0.05 m with band levels N(80, 6) dB plus the A-weighting shape; 500 rotated
  rectangles (sides U[5, 40] m, heights U[3, 30] m); p_fav = 0.5; ISO 9613-1 air at 15 C / 70 %.
- ``tile_scene`` (configuration "plain exact painter at tile scale"): receivers at 10 m over a
  diamond-square DEM (30 m lattice, 200 m relief), point sources (road 0.05 m / rail 0.5 m) with
  band levels N(85, 8) dB and day/evening/night offsets, polygon buildings with 4-12 edges and
  heights U[3, 60] m, 2 km propagation cutoff. Its full size is 4096x4096 receivers, 200,000
  sources and 50,000 buildings; ``scale`` shrinks the area (and the counts with it) for tests.

The generators use numpy's PCG64 with the configuration's seed; they are not the reference's
SplitMix64 streams (the reference publishes no such scenes), so the scenes are this project's
own, fixed by seed.
"""
from __future__ import annotations

import math

import numpy as np

from . import abi

A_WEIGHTING_DB = np.array([-26.2, -16.1, -8.6, -3.2, 0.0, 1.2, 1.0, -1.1])
# ISO 9613-1 at 15 C, 70 %, 101.325 kPa per octave band (dB/km), as physics/src/atmosphere.rs
# computes it; coded to 0.01 dB/km in the weather node as the reference's table stores it.
ALPHA_15C_70 = np.array([0.104927, 0.380965, 1.131497, 2.362992, 4.07924, 8.748437, 26.385662, 93.713703])
METRES_PER_STEP = 0.2


def _weather(origin, extent, percent_of, alpha=ALPHA_15C_70, spacing=1000.0):
    columns = int(math.ceil(extent[0] / spacing)) + 2
    rows = int(math.ceil(extent[1] / spacing)) + 2
    nodes = np.zeros((rows, columns, abi.WEATHER_NODE_BYTES), np.uint8)
    for r in range(rows):
        for c in range(columns):
            nodes[r, c] = abi.Weather.node(percent_of(r, c), alpha)
    return abi.Weather(origin[0] - spacing / 2, origin[1] - spacing / 2, spacing, nodes)


def _obstacles(rings, heights, kinds, footprints, origin_m, extent_m, metres_per_step=METRES_PER_STEP,
               metres_per_step_y=None):
    """The cell index of outlines given in metres: vertices snapped to integer steps (y south), each
    ring closed by repeating its first vertex, every edge listed in every cell its step box
    (padded by one step) touches, consecutive edges of one outline in one cell joined as a run.
    ``metres_per_step_y``: north metres per step where they differ from the east's."""
    mpx = metres_per_step
    mpy = metres_per_step if metres_per_step_y is None else metres_per_step_y
    origin_steps = (10_000.0, 10_000.0 + extent_m[1] / mpy + 5_000.0)
    cs = abi.CELL_STEPS
    # Grid covering the area plus a margin of cells.
    low_steps = (origin_steps[0] + (origin_m[0] - 2000.0) / mpx,
                 origin_steps[1] - (origin_m[1] + extent_m[1] + 2000.0) / mpy)
    high_steps = (origin_steps[0] + (origin_m[0] + extent_m[0] + 2000.0) / mpx,
                  origin_steps[1] - (origin_m[1] - 2000.0) / mpy)
    cell_x0, cell_y0 = int(math.floor(low_steps[0] / cs)), int(math.floor(low_steps[1] / cs))
    cells_x = int(math.floor(high_steps[0] / cs)) - cell_x0 + 1
    cells_y = int(math.floor(high_steps[1] / cs)) - cell_y0 + 1
    vertices = []
    outlines = np.zeros(len(rings), abi.OUTLINE_DTYPE)
    per_cell = {}
    first = 0
    for o, ring in enumerate(rings):
        ring = np.asarray(ring, np.float64)
        sx = np.round(origin_steps[0] + ring[:, 0] / mpx).astype(np.int64)
        sy = np.round(origin_steps[1] - ring[:, 1] / mpy).astype(np.int64)
        if kinds[o] == abi.KIND_BUILDING:
            sx, sy = np.append(sx, sx[0]), np.append(sy, sy[0])
        pts = np.stack([sx, sy], 1)
        vertices.append(pts)
        outlines[o]["footprint_id"] = footprints[o]
        outlines[o]["first_vertex"] = first
        outlines[o]["vertex_count"] = len(pts)
        outlines[o]["height_dm"] = int(round(heights[o] * 10.0))
        outlines[o]["kind"] = kinds[o]
        for e in range(len(pts) - 1):
            (ax, ay), (bx, by) = pts[e], pts[e + 1]
            c0 = (min(ax, bx) - 1) // cs - cell_x0, (min(ay, by) - 1) // cs - cell_y0
            c1 = (max(ax, bx) + 1) // cs - cell_x0, (max(ay, by) + 1) // cs - cell_y0
            for cy in range(c0[1], c1[1] + 1):
                for cx in range(c0[0], c1[0] + 1):
                    per_cell.setdefault(cy * cells_x + cx, []).append((o, e))
        first += len(pts)
    runs = []
    cell_first_run = np.zeros(cells_x * cells_y + 1, np.uint32)
    for index in range(cells_x * cells_y):
        cell_first_run[index] = len(runs)
        edges = per_cell.get(index)
        if not edges:
            continue
        edges.sort()
        start = edges[0]
        count = 1
        for prev, cur in zip(edges, edges[1:]):
            if cur[0] == prev[0] and cur[1] == prev[1] + 1:
                count += 1
            else:
                runs.append((start[0], start[1], count))
                start, count = cur, 1
        runs.append((start[0], start[1], count))
    cell_first_run[-1] = len(runs)
    runs_arr = np.zeros(len(runs), abi.RUN_DTYPE)
    if runs:
        r = np.array(runs, np.int64)
        runs_arr["outline"], runs_arr["first_edge"], runs_arr["edge_count"] = r[:, 0], r[:, 1], r[:, 2]
    verts = np.concatenate(vertices) if vertices else np.zeros((0, 2), np.int64)
    assert verts.size == 0 or (np.abs(verts).max() < 2**31)
    return abi.Obstacles(origin_steps, mpx, (cell_x0, cell_y0), (cells_x, cells_y), cell_first_run,
                         runs_arr, outlines, verts.astype(np.int32), metres_per_step_y=0.0 if metres_per_step_y is None
                         else mpy)


def _inside(ring, x, y):
    """Even-odd point-in-polygon of points (x, y) against one ring (metres)."""
    inside = np.zeros(x.shape, bool)
    n = len(ring)
    for i in range(n):
        ax, ay = ring[i]
        bx, by = ring[(i + 1) % n]
        crosses = (ay > y) != (by > y)
        with np.errstate(divide="ignore", invalid="ignore"):
            xi = ax + (y - ay) * (bx - ax) / (by - ay)
        inside ^= crosses & (x < xi)
    return inside


def _own_footprints(points, rings, footprints, spacing, origin, shape):
    """A receiver inside a building stands in it (paint/src/exact.rs:36-41): its footprint id."""
    nx, ny = shape
    own = points["own_footprint"].reshape(ny, nx)
    for ring, fid in zip(rings, footprints):
        ring = np.asarray(ring)
        i0 = max(int(math.floor((ring[:, 0].min() - origin[0]) / spacing)), 0)
        i1 = min(int(math.ceil((ring[:, 0].max() - origin[0]) / spacing)), nx - 1)
        j0 = max(int(math.floor((ring[:, 1].min() - origin[1]) / spacing)), 0)
        j1 = min(int(math.ceil((ring[:, 1].max() - origin[1]) / spacing)), ny - 1)
        if i1 < i0 or j1 < j0:
            continue
        jj, ii = np.mgrid[j0 : j1 + 1, i0 : i1 + 1]
        x, y = origin[0] + ii * spacing, origin[1] + jj * spacing
        hit = _inside(ring, x, y) & (own[j0 : j1 + 1, i0 : i1 + 1] == 0)
        own[j0 : j1 + 1, i0 : i1 + 1][hit] = fid
    points["own_footprint"] = own.reshape(-1)


def receiver_grid(nx, ny, spacing, origin):
    """Receivers at the centres of a grid of cells (no receiver on a lattice line)."""
    points = np.zeros(nx * ny, abi.POINT_DTYPE)
    jj, ii = np.mgrid[0:ny, 0:nx]
    points["x_m"] = origin[0] + ii.reshape(-1) * spacing
    points["y_m"] = origin[1] + jj.reshape(-1) * spacing
    return points


def _polyline_points(rng, count_lines, per_line, extent, origin):
    out = []
    for _ in range(count_lines):
        k = rng.integers(2, 6)
        verts = np.stack([origin[0] + rng.uniform(0, extent[0], k), origin[1] + rng.uniform(0, extent[1], k)], 1)
        seg = np.linalg.norm(np.diff(verts, axis=0), axis=1)
        total = seg.sum()
        cum = np.concatenate([[0.0], np.cumsum(seg)])
        for s in (np.arange(per_line) + 0.5) / per_line * total:
            i = min(int(np.searchsorted(cum, s, side="right")) - 1, k - 2)
            f = (s - cum[i]) / seg[i] if seg[i] > 0 else 0.0
            out.append(verts[i] + f * (verts[i + 1] - verts[i]))
    return np.array(out)


class SyntheticScene:
    """A scene with its sources, receivers and cutoff."""

    def __init__(self, name, scene, sources, points, grid_shape, cutoff_m, rings=None):
        self.name, self.scene, self.sources, self.points = name, scene, sources, points
        self.grid_shape, self.cutoff_m, self.rings = grid_shape, cutoff_m, rings

    def window(self, side):
        """The central side x side receivers (row-major) and the window's shape (nx, ny)."""
        gx, gy = self.grid_shape
        nx, ny = min(side, gx), min(side, gy)
        x0, y0 = (gx - nx) // 2, (gy - ny) // 2
        index = ((np.arange(ny)[:, None] + y0) * gx + (np.arange(nx)[None, :] + x0)).reshape(-1)
        return np.ascontiguousarray(self.points[index]), nx, ny


def road_scene(seed=1, receivers=256, spacing=10.0, lines=200, per_line=10, buildings=500):
    """Configuration "plain CNOSSOS road scene": see the module docstring."""
    rng = np.random.default_rng(seed)
    extent = receivers * spacing
    origin = (0.0, 0.0)
    # Terrain: 10 m lattice from -200 m to extent + 200 m, flat at 0 m, G per node {0, 0.5, 1}.
    n = int(round((extent + 400.0) / 10.0)) + 1
    impervious = rng.choice(np.array([100, 50, 0]), size=(n, n))
    terrain = abi.Terrain(-200.0, -200.0, 10.0, abi.Terrain.pack(np.zeros((n, n)), impervious))
    # Buildings: rotated rectangles.
    centres = rng.uniform(0, extent, (buildings, 2))
    sides = rng.uniform(5.0, 40.0, (buildings, 2))
    angles = rng.uniform(0, math.pi, buildings)
    heights = rng.uniform(3.0, 30.0, buildings)
    rings = []
    for (cx, cy), (a, b), th in zip(centres, sides, angles):
        c, s = math.cos(th), math.sin(th)
        corners = [(-a / 2, -b / 2), (a / 2, -b / 2), (a / 2, b / 2), (-a / 2, b / 2)]
        rings.append([(cx + c * u - s * v, cy + s * u + c * v) for u, v in corners])
    footprints = np.arange(1, buildings + 1, dtype=np.uint64)
    obstacles = _obstacles(rings, heights, [abi.KIND_BUILDING] * buildings, footprints, origin, (extent, extent))
    weather = _weather((-1000.0, -1000.0), (extent + 2000.0, extent + 2000.0),
                       lambda r, c: np.full((3, 16), 50))
    # Sources: 200 polylines x 10 points at 0.05 m, band levels N(80, 6) + A-weighting shape.
    xy = _polyline_points(rng, lines, per_line, (extent, extent), origin)
    sources = np.zeros(len(xy), abi.SOURCE_DTYPE)
    sources["x_m"], sources["y_m"] = xy[:, 0], xy[:, 1]
    sources["height_m"] = 0.05
    sources["exclusion_radius_m"] = 2.0
    sources["ground_percent"] = abi.GROUND_FROM_TERRAIN
    levels = rng.normal(80.0, 6.0, (len(xy), 8)) + A_WEIGHTING_DB
    energy = 10.0 ** (levels / 10.0)
    sources["energy"] = np.repeat(energy[:, None, :], 3, axis=1)
    points = receiver_grid(receivers, receivers, spacing, (spacing / 2, spacing / 2))
    _own_footprints(points, rings, footprints, spacing, (spacing / 2, spacing / 2), (receivers, receivers))
    scene = abi.Scene(terrain, obstacles, weather)
    return SyntheticScene("road", scene, sources, points, (receivers, receivers), float("inf"), rings)


# The geographic frame (tiles/src/geo.rs): z12 Web Mercator units, WGS84 metres per unit at the
# frame's origin, the one arc-second terrain lattice and the world weather table.
TILES_PER_AXIS = 4096
WGS84_A_M = 6_378_137.0
WGS84_E2 = 6.694_379_990_141_316e-3
TERRAIN_NODES_PER_DEGREE = 3600
STEPS_PER_TILE = 32_768.0


def mercator(lat_deg, lon_deg):
    """geo::Mercator::from_degrees: (x, y) in z12 tile units, y growing south."""
    n = float(TILES_PER_AXIS)
    lon = (np.asarray(lon_deg, np.float64) + 180.0) % 360.0 - 180.0
    phi = np.radians(np.asarray(lat_deg, np.float64))
    return (lon + 180.0) / 360.0 * n, (1.0 - np.arcsinh(np.tan(phi)) / math.pi) / 2.0 * n


def mercator_degrees(x, y):
    """geo::Mercator::to_degrees: (latitude, longitude)."""
    n = float(TILES_PER_AXIS)
    lat = np.degrees(np.arctan(np.sinh(math.pi * (1.0 - 2.0 * np.asarray(y, np.float64) / n))))
    return lat, np.asarray(x, np.float64) / n * 360.0 - 180.0


def local_frame(origin_x, origin_y):
    """geo::LocalFrame::at: east and north metres per Mercator unit at the origin (WGS84)."""
    lat, _ = mercator_degrees(origin_x, origin_y)
    phi = math.radians(float(lat))
    w = 1.0 - WGS84_E2 * math.sin(phi) ** 2
    radians_per_unit = 2.0 * math.pi / TILES_PER_AXIS
    east = WGS84_A_M / math.sqrt(w) * math.cos(phi) * radians_per_unit
    north = WGS84_A_M * (1.0 - WGS84_E2) / w ** 1.5 * math.cos(phi) * radians_per_unit
    return east, north


def geo_scene(seed=21, lat_deg=50.0875, lon_deg=14.4214, receivers=64, spacing=10.0, lines=30, per_line=10,
              buildings=150, relief_m=40.0, npd=2):
    """A square in the geographic frame: a frame at (lat, lon), the one arc-second terrain lattice
    about the painted area (rows at their Mercator latitudes: ``row_y_m``; columns evenly:
    ``column_m``), smooth terrain of `relief_m`, buildings and road points as in the road scene, the
    obstacle steps' east and north metres from the frame (STEPS_PER_TILE a unit) and the world
    weather table (`npd` nodes a degree) read by latitude and longitude."""
    rng = np.random.default_rng(seed)
    ox, oy = mercator(lat_deg, lon_deg)
    east_m, north_m = local_frame(ox, oy)
    extent = receivers * spacing
    margin = 400.0
    # The lattice's rows and columns about [-margin, extent + margin] in both axes.
    lat_lo, _ = mercator_degrees(ox, oy + margin / north_m)
    lat_hi, _ = mercator_degrees(ox, oy - (extent + margin) / north_m)
    _, lon_lo = mercator_degrees(ox - margin / east_m, oy)
    _, lon_hi = mercator_degrees(ox + (extent + margin) / east_m, oy)
    row_lo, row_hi = int(math.floor(lat_lo * TERRAIN_NODES_PER_DEGREE)), int(math.ceil(lat_hi * TERRAIN_NODES_PER_DEGREE))
    col_lo, col_hi = int(math.floor(lon_lo * TERRAIN_NODES_PER_DEGREE)), int(math.ceil(lon_hi * TERRAIN_NODES_PER_DEGREE))
    row_lat = np.arange(row_lo, row_hi + 1) / TERRAIN_NODES_PER_DEGREE
    _, row_y = mercator(row_lat, np.zeros_like(row_lat))
    row_y_m = (oy - row_y) * north_m
    col_lon = np.arange(col_lo, col_hi + 1) / TERRAIN_NODES_PER_DEGREE
    col_x, _ = mercator(np.zeros_like(col_lon), col_lon)
    col_east = (col_x - ox) * east_m
    column_m = float(col_east[1] - col_east[0])
    ee, nn = np.meshgrid(col_east, row_y_m)
    height = 200.0 + relief_m * (0.5 * np.sin(ee / 310.0) * np.cos(nn / 260.0) + 0.3 * np.sin((ee + 2 * nn) / 170.0))
    height += rng.normal(0.0, 0.3, height.shape)
    impervious = rng.choice(np.array([100, 50, 0]), size=height.shape)
    terrain = abi.Terrain(float(col_east[0]), float(row_y_m[0]), float(np.mean(np.diff(row_y_m))),
                          abi.Terrain.pack(height, impervious), column_m=column_m, row_y_m=row_y_m)
    # Buildings: rotated rectangles; obstacle steps of the frame (anisotropic).
    centres = rng.uniform(0, extent, (buildings, 2))
    sides = rng.uniform(5.0, 40.0, (buildings, 2))
    angles = rng.uniform(0, math.pi, buildings)
    heights = rng.uniform(3.0, 30.0, buildings)
    rings = []
    for (cx, cy), (a, b), th in zip(centres, sides, angles):
        c, s_ = math.cos(th), math.sin(th)
        corners = [(-a / 2, -b / 2), (a / 2, -b / 2), (a / 2, b / 2), (-a / 2, b / 2)]
        rings.append([(cx + c * u - s_ * v, cy + s_ * u + c * v) for u, v in corners])
    footprints = np.arange(1, buildings + 1, dtype=np.uint64)
    obstacles = _obstacles(rings, heights, [abi.KIND_BUILDING] * buildings, footprints, (0.0, 0.0), (extent, extent),
                           metres_per_step=east_m / STEPS_PER_TILE, metres_per_step_y=north_m / STEPS_PER_TILE)
    # The world weather table: favourable percentages and absorptions varying node by node.
    rows_w, cols_w = 180 * npd + 1, 360 * npd
    nodes = np.zeros((rows_w, cols_w, abi.WEATHER_NODE_BYTES), np.uint8)
    rr, cc = np.meshgrid(np.arange(rows_w), np.arange(cols_w), indexing="ij")
    for p in range(3):
        for k in range(16):
            nodes[:, :, p * 16 + k] = (20 + (rr * 7 + cc * 13 + p * 11 + k * 5) % 61).astype(np.uint8)
    alpha = np.round(ALPHA_15C_70 * 100.0)[None, None, :] * (1.0 + 0.02 * ((rr + cc) % 7))[:, :, None]
    nodes[:, :, 48:] = alpha.astype("<u2").view(np.uint8).reshape(rows_w, cols_w, 16)
    weather = abi.Weather(0.0, 0.0, 1.0, nodes, frame=(ox, oy, east_m, north_m), nodes_per_degree=npd)
    xy = _polyline_points(rng, lines, per_line, (extent, extent), (0.0, 0.0))
    sources = np.zeros(len(xy), abi.SOURCE_DTYPE)
    sources["x_m"], sources["y_m"] = xy[:, 0], xy[:, 1]
    sources["height_m"] = 0.05
    sources["exclusion_radius_m"] = 2.0
    sources["ground_percent"] = abi.GROUND_FROM_TERRAIN
    levels = rng.normal(80.0, 6.0, (len(xy), 8)) + A_WEIGHTING_DB
    sources["energy"] = np.repeat((10.0 ** (levels / 10.0))[:, None, :], 3, axis=1)
    points = receiver_grid(receivers, receivers, spacing, (spacing / 2, spacing / 2))
    _own_footprints(points, rings, footprints, spacing, (spacing / 2, spacing / 2), (receivers, receivers))
    scene = abi.Scene(terrain, obstacles, weather)
    sc = SyntheticScene("geo", scene, sources, points, (receivers, receivers), float("inf"), rings)
    sc.frame = (ox, oy, east_m, north_m)
    return sc


def _diamond_square(rng, power, relief):
    n = 2**power + 1
    h = np.zeros((n, n))
    h[0, 0], h[0, -1], h[-1, 0], h[-1, -1] = rng.uniform(0, 1, 4)
    step, scale = n - 1, 1.0
    while step > 1:
        half = step // 2
        # Diamond step.
        c = (h[0:-1:step, 0:-1:step] + h[0:-1:step, step::step] + h[step::step, 0:-1:step] + h[step::step, step::step])
        h[half::step, half::step] = c / 4.0 + rng.uniform(-scale, scale, c.shape)
        # Square step.
        for (oy, ox) in ((0, half), (half, 0)):
            ys, xs = np.mgrid[oy:n:step, ox:n:step]
            total = np.zeros(ys.shape)
            count = np.zeros(ys.shape)
            for dy, dx in ((-half, 0), (half, 0), (0, -half), (0, half)):
                yy, xx = ys + dy, xs + dx
                ok = (yy >= 0) & (yy < n) & (xx >= 0) & (xx < n)
                total[ok] += h[yy[ok], xx[ok]]
                count[ok] += 1
            h[ys, xs] = total / count + rng.uniform(-scale, scale, ys.shape)
        step, scale = half, scale * 0.55
    h -= h.min()
    return h * (relief / h.max())


def tile_scene(seed=2, scale=1.0, receivers=4096, spacing=10.0, sources=200_000, buildings=50_000, cutoff_m=2000.0):
    """Configuration "plain exact painter at tile scale"; ``scale`` < 1 shrinks the square side by
    that factor and the source and building counts by its square (same densities)."""
    rng = np.random.default_rng(seed)
    nrec = max(int(round(receivers * scale)), 8)
    extent = nrec * spacing
    area_ratio = (nrec / receivers) ** 2
    n_sources = max(int(round(sources * area_ratio)), 16)
    n_buildings = max(int(round(buildings * area_ratio)), 8)
    origin = (0.0, 0.0)
    # DEM: diamond-square, 30 m lattice, 200 m relief, covering the area and 600 m around it.
    margin = 600.0
    nodes_needed = int(math.ceil((extent + 2 * margin) / 30.0)) + 1
    power = max(int(math.ceil(math.log2(max(nodes_needed - 1, 2)))), 2)
    dem = _diamond_square(rng, power, 200.0)[:nodes_needed, :nodes_needed] + 100.0
    impervious = np.clip(rng.normal(40, 30, dem.shape), 0, 100).astype(np.uint32)
    terrain = abi.Terrain(-margin, -margin, 30.0, abi.Terrain.pack(dem, impervious))
    # Buildings: star-shaped polygons of 4-12 edges, radii 3-20 m, heights U[3, 60] m.
    centres = rng.uniform(0, extent, (n_buildings, 2))
    heights = rng.uniform(3.0, 60.0, n_buildings)
    rings = []
    for cx, cy in centres:
        k = int(rng.integers(4, 13))
        ang = np.sort(rng.uniform(0, 2 * math.pi, k))
        base = rng.uniform(5.0, 20.0)
        rad = base * rng.uniform(0.6, 1.0, k)
        rings.append(list(zip(cx + rad * np.cos(ang), cy + rad * np.sin(ang))))
    footprints = np.arange(1, n_buildings + 1, dtype=np.uint64)
    obstacles = _obstacles(rings, heights, [abi.KIND_BUILDING] * n_buildings, footprints, origin, (extent, extent))
    weather = _weather((-margin, -margin), (extent + 2 * margin, extent + 2 * margin),
                       lambda r, c: rng.integers(20, 80, (3, 16)))
    # Sources: road (0.05 m) and rail (0.5 m) points along random polylines.
    per_line = 10
    xy = _polyline_points(rng, int(math.ceil(n_sources / per_line)), per_line, (extent, extent), origin)[:n_sources]
    src = np.zeros(len(xy), abi.SOURCE_DTYPE)
    src["x_m"], src["y_m"] = xy[:, 0], xy[:, 1]
    rail = rng.uniform(0, 1, len(xy)) < 0.3
    src["height_m"] = np.where(rail, 0.5, 0.05)
    src["layer"] = np.where(rail, 1, 0)
    src["exclusion_radius_m"] = 2.0
    src["ground_percent"] = abi.GROUND_FROM_TERRAIN
    levels = rng.normal(85.0, 8.0, (len(xy), 8)) + A_WEIGHTING_DB
    period_offset = np.array([0.0, -2.0, -7.0])
    src["energy"] = 10.0 ** ((levels[:, None, :] + period_offset[None, :, None]) / 10.0)
    points = receiver_grid(nrec, nrec, spacing, (spacing / 2, spacing / 2))
    _own_footprints(points, rings, footprints, spacing, (spacing / 2, spacing / 2), (nrec, nrec))
    scene = abi.Scene(terrain, obstacles, weather)
    return SyntheticScene("tile", scene, src, points, (nrec, nrec), cutoff_m, rings)


class AircraftScene:
    """Receivers, noise classes, tracks and their vertices of the aircraft configuration."""

    def __init__(self, points_xy, grid_shape, classes, tracks, vertices, receiver_altitude_m, reach_m):
        self.points_xy = points_xy
        self.grid_shape = grid_shape
        self.classes, self.tracks, self.vertices = classes, tracks, vertices
        self.receiver_altitude_m, self.reach_m = receiver_altitude_m, reach_m


def aircraft_scene(seed=3, scale=1.0, receivers=2048, spacing=50.0, classes=100, tracks=20_000, segments=200,
                   segment_m=250.0, reach_m=16_000.0):
    """Configuration "Doc 29 aircraft": 100 noise classes with NPD tables (10 distances x 4 power
    settings per operation, SEL at 1,000 ft ~ N(90, 8) dB), 20,000 tracks of 200 segments with 3 degree
    climbs and descents from four airports, a 2048x2048 receiver grid at 50 m (4 m above the datum),
    16 km reach. ``scale`` shrinks the grid side and the track count with the area."""
    rng = np.random.default_rng(seed)
    n = max(int(round(receivers * scale)), 4)
    extent = n * spacing
    n_tracks = max(int(round(tracks * (n / receivers) ** 2)), 8)
    d_ft = np.array([200.0, 400.0, 630.0, 1000.0, 2000.0, 4000.0, 6300.0, 10000.0, 16000.0, 25000.0])
    cl = np.zeros(classes, abi.AIRCRAFT_CLASS_DTYPE)
    for k in range(classes):
        u = rng.uniform()
        cl[k]["installation"] = abi.INSTALL_WING if u < 0.5 else (abi.INSTALL_FUSELAGE if u < 0.8 else abi.INSTALL_PROPELLER)
        cl[k]["v_ref_kt"] = rng.uniform(140.0, 180.0)
        placeholder = rng.uniform() < 0.05
        for dep in range(2):
            cl[k]["rows"][dep] = abi.POWER_ROWS
            base = np.clip(rng.normal(90.0, 8.0), 65.0, 115.0) + (3.0 if dep else 0.0)
            slope = rng.uniform(13.0, 19.0)
            absorb = rng.uniform(0.0006, 0.0012)
            for r in range(abi.POWER_ROWS):
                sel = base + 2.0 * r - slope * np.log10(d_ft / 1000.0) - absorb * (d_ft - 1000.0)
                delta = 12.0 if placeholder else 8.0 + 6.0 * np.log10(d_ft / 1000.0) + rng.uniform(-1.0, 1.0)
                cl[k]["sel_db"][dep, r] = sel
                cl[k]["lamax_db"][dep, r] = sel - delta
    airports = rng.uniform(0.2 * extent, 0.8 * extent, (4, 2))
    headings = rng.uniform(0.0, np.pi, 4)
    tr = np.zeros(n_tracks, abi.TRACK_DTYPE)
    verts = np.zeros((n_tracks, segments + 1, 3), np.float32)
    tan3 = np.tan(np.radians(3.0))
    for t in range(n_tracks):
        a = rng.integers(0, 4)
        dep = int(rng.uniform() < 0.5)
        heading = headings[a] + (np.pi if rng.uniform() < 0.5 else 0.0)
        turn = rng.normal(0.0, 0.004, segments)
        turn[: segments // 10] = 0.0
        head = heading + np.cumsum(turn)
        steps = np.stack([np.cos(head), np.sin(head)], 1) * segment_m
        path = np.concatenate([[airports[a]], airports[a] + np.cumsum(steps, 0)])
        along = np.arange(segments + 1) * segment_m
        z = np.minimum(along * tan3, 3000.0)
        if not dep:  # an approach: flown toward the threshold, descending
            path, z = path[::-1], z[::-1]
        verts[t, :, 0], verts[t, :, 1], verts[t, :, 2] = path[:, 0], path[:, 1], z
        tr[t]["first_vertex"] = t * (segments + 1)
        tr[t]["vertex_count"] = segments + 1
        tr[t]["aircraft_class"] = rng.integers(0, classes)
        tr[t]["departure"] = dep
        if dep:
            tr[t]["power_row"], tr[t]["power_weight"] = 2, rng.uniform(0.0, 1.0)
        else:
            tr[t]["power_row"], tr[t]["power_weight"] = rng.integers(0, 2), rng.uniform(0.0, 1.0)
        tr[t]["speed_kt"] = rng.uniform(140.0, 220.0)
        tr[t]["flights"] = [rng.poisson(3.0), rng.poisson(1.0), rng.poisson(0.5)]
    jj, ii = np.mgrid[0:n, 0:n]
    xy = np.stack([(ii.reshape(-1) + 0.5) * spacing, (jj.reshape(-1) + 0.5) * spacing], 1).astype(np.float64)
    return AircraftScene(np.ascontiguousarray(xy), (n, n), cl, tr, verts.reshape(-1, 3), 4.0, reach_m)


class TileMosaic:
    """Tiles of receivers, each painted against its own sources only (the batched configuration).

    `scene` covers the whole mosaic; `points` are tile-major (tile t's receivers are
    points[t*nx*ny : (t+1)*nx*ny], row-major); `sources` are sorted by tile, tile t's being
    sources[source_first[t] : source_first[t+1]]."""

    def __init__(self, scene, sources, points, tile_shape, tiles_xy, source_first):
        self.scene, self.sources, self.points = scene, sources, points
        self.tile_shape, self.tiles_xy, self.source_first = tile_shape, tiles_xy, source_first

    @property
    def n_tiles(self):
        return self.tiles_xy[0] * self.tiles_xy[1]

    def tile_table(self):
        """The qm_tile records of every tile (abi.TILE_DTYPE)."""
        nx, ny = self.tile_shape
        t = np.zeros(self.n_tiles, abi.TILE_DTYPE)
        t["receiver_first"] = np.arange(self.n_tiles, dtype=np.uint32) * nx * ny
        t["nx"], t["ny"] = nx, ny
        t["source_first"] = self.source_first[:-1]
        t["source_count"] = np.diff(self.source_first)
        return t


def tile_mosaic(seed=4, tiles=(4, 4), receivers=256, spacing=10.0, source_median=500, building_median=300,
                sigma=1.0, source_clip=(10, 50_000), building_clip=(1, 50_000)):
    """Configuration "batched uneven tiles": tiles of receivers x receivers at `spacing`, per tile a
    log-normal count of point sources (median 500, sigma 1, clipped to 10..50k) and of buildings
    (median 300), over one diamond-square DEM (30 m lattice, 200 m relief)."""
    rng = np.random.default_rng(seed)
    tx, ty = tiles
    side = receivers * spacing
    extent = (tx * side, ty * side)
    margin = 600.0
    nodes_needed = int(math.ceil((max(extent) + 2 * margin) / 30.0)) + 1
    power = max(int(math.ceil(math.log2(max(nodes_needed - 1, 2)))), 2)
    dem = _diamond_square(rng, power, 200.0)[:nodes_needed, :nodes_needed] + 100.0
    impervious = np.clip(rng.normal(40, 30, dem.shape), 0, 100).astype(np.uint32)
    terrain = abi.Terrain(-margin, -margin, 30.0, abi.Terrain.pack(dem, impervious))
    n_tiles = tx * ty
    n_src = np.clip(np.round(source_median * np.exp(sigma * rng.normal(size=n_tiles))), *source_clip).astype(int)
    n_bld = np.clip(np.round(building_median * np.exp(sigma * rng.normal(size=n_tiles))), *building_clip).astype(int)
    rings, heights = [], []
    all_src, first = [], [0]
    for t in range(n_tiles):
        ox, oy = (t % tx) * side, (t // tx) * side
        for cx, cy in rng.uniform(0, side, (n_bld[t], 2)) + (ox, oy):
            k = int(rng.integers(4, 13))
            ang = np.sort(rng.uniform(0, 2 * math.pi, k))
            rad = rng.uniform(5.0, 20.0) * rng.uniform(0.6, 1.0, k)
            rings.append(list(zip(cx + rad * np.cos(ang), cy + rad * np.sin(ang))))
            heights.append(rng.uniform(3.0, 60.0))
        per_line = 10
        xy = _polyline_points(rng, int(math.ceil(n_src[t] / per_line)), per_line, (side, side), (ox, oy))[:n_src[t]]
        src = np.zeros(len(xy), abi.SOURCE_DTYPE)
        src["x_m"], src["y_m"] = xy[:, 0], xy[:, 1]
        rail = rng.uniform(0, 1, len(xy)) < 0.3
        src["height_m"] = np.where(rail, 0.5, 0.05)
        src["layer"] = np.where(rail, 1, 0)
        src["exclusion_radius_m"] = 2.0
        src["ground_percent"] = abi.GROUND_FROM_TERRAIN
        levels = rng.normal(85.0, 8.0, (len(xy), 8)) + A_WEIGHTING_DB
        period_offset = np.array([0.0, -2.0, -7.0])
        src["energy"] = 10.0 ** ((levels[:, None, :] + period_offset[None, :, None]) / 10.0)
        all_src.append(src)
        first.append(first[-1] + len(src))
    footprints = np.arange(1, len(rings) + 1, dtype=np.uint64)
    obstacles = _obstacles(rings, heights, [abi.KIND_BUILDING] * len(rings), footprints, (0.0, 0.0), extent)
    weather = _weather((-margin, -margin), (extent[0] + 2 * margin, extent[1] + 2 * margin),
                       lambda r, c: rng.integers(20, 80, (3, 16)))
    pts = []
    for t in range(n_tiles):
        ox, oy = (t % tx) * side, (t // tx) * side
        p = receiver_grid(receivers, receivers, spacing, (ox + spacing / 2, oy + spacing / 2))
        _own_footprints(p, rings, footprints, spacing, (ox + spacing / 2, oy + spacing / 2), (receivers, receivers))
        pts.append(p)
    scene = abi.Scene(terrain, obstacles, weather)
    return TileMosaic(scene, np.concatenate(all_src), np.concatenate(pts), (receivers, receivers), (tx, ty),
                      np.array(first, np.uint32))


def _obstacles_fast(rings_xy, ring_size, heights, origin_m, extent_m, metres_per_step=METRES_PER_STEP):
    """The cell index of many closed polygons at once (as _obstacles, vectorized): `rings_xy` the
    vertices (total, 2) in metres of rings of `ring_size` vertices each (not yet closed). Each
    edge is listed in every cell of its step box padded by one step; an outline's consecutive
    edges in one cell form one run."""
    mps = metres_per_step
    origin_steps = (10_000.0, 10_000.0 + extent_m[1] / mps + 5_000.0)
    cs = abi.CELL_STEPS
    cell_x0 = int(math.floor((origin_steps[0] + (origin_m[0] - 2000.0) / mps) / cs))
    cell_y0 = int(math.floor((origin_steps[1] - (origin_m[1] + extent_m[1] + 2000.0) / mps) / cs))
    cells_x = int(math.floor((origin_steps[0] + (origin_m[0] + extent_m[0] + 2000.0) / mps) / cs)) - cell_x0 + 1
    cells_y = int(math.floor((origin_steps[1] - (origin_m[1] - 2000.0) / mps) / cs)) - cell_y0 + 1
    ring_size = np.asarray(ring_size, np.int64)
    n_out = ring_size.size
    first_in = np.concatenate([[0], np.cumsum(ring_size)[:-1]])
    sx = np.round(origin_steps[0] + rings_xy[:, 0] / mps).astype(np.int64)
    sy = np.round(origin_steps[1] - rings_xy[:, 1] / mps).astype(np.int64)
    # Closed rings: each outline's first vertex repeated after its last.
    closed = ring_size + 1
    first_out = np.concatenate([[0], np.cumsum(closed)[:-1]])
    total = int(closed.sum())
    vx, vy = np.empty(total, np.int64), np.empty(total, np.int64)
    owner_in = np.repeat(np.arange(n_out), ring_size)
    pos_in = np.arange(sx.size) - first_in[owner_in]
    vx[first_out[owner_in] + pos_in], vy[first_out[owner_in] + pos_in] = sx, sy
    vx[first_out + ring_size], vy[first_out + ring_size] = sx[first_in], sy[first_in]
    # Edges: outline o, edge e from vertex e to e + 1.
    e_owner = np.repeat(np.arange(n_out), ring_size)
    e_index = np.arange(e_owner.size) - np.concatenate([[0], np.cumsum(ring_size)[:-1]])[e_owner]
    a = first_out[e_owner] + e_index
    ax, ay, bx, by = vx[a], vy[a], vx[a + 1], vy[a + 1]
    c0x = (np.minimum(ax, bx) - 1) // cs - cell_x0
    c0y = (np.minimum(ay, by) - 1) // cs - cell_y0
    c1x = (np.maximum(ax, bx) + 1) // cs - cell_x0
    c1y = (np.maximum(ay, by) + 1) // cs - cell_y0
    wx, wy = c1x - c0x + 1, c1y - c0y + 1
    count = wx * wy
    k = np.repeat(np.arange(e_owner.size), count)
    local = np.arange(k.size) - np.concatenate([[0], np.cumsum(count)[:-1]])[k]
    cell = (c0y[k] + local // wx[k]) * cells_x + (c0x[k] + local % wx[k])
    order = np.lexsort((e_index[k], e_owner[k], cell))
    cell, o, e = cell[order], e_owner[k][order], e_index[k][order]
    new_run = np.ones(cell.size, bool)
    new_run[1:] = (cell[1:] != cell[:-1]) | (o[1:] != o[:-1]) | (e[1:] != e[:-1] + 1)
    starts = np.flatnonzero(new_run)
    lengths = np.diff(np.append(starts, cell.size))
    runs = np.zeros(starts.size, abi.RUN_DTYPE)
    runs["outline"], runs["first_edge"], runs["edge_count"] = o[starts], e[starts], lengths
    run_cell = cell[starts]
    cell_first_run = np.searchsorted(run_cell, np.arange(cells_x * cells_y + 1)).astype(np.uint32)
    outlines = np.zeros(n_out, abi.OUTLINE_DTYPE)
    outlines["footprint_id"] = np.arange(1, n_out + 1, dtype=np.uint64)
    outlines["first_vertex"] = first_out
    outlines["vertex_count"] = closed
    outlines["height_dm"] = np.round(np.asarray(heights) * 10.0).astype(np.uint16)
    outlines["kind"] = abi.KIND_BUILDING
    verts = np.stack([vx, vy], 1).astype(np.int32)
    return abi.Obstacles(origin_steps, mps, (cell_x0, cell_y0), (cells_x, cells_y), cell_first_run, runs, outlines,
                         verts)


def tile_mosaic_fast(seed=4, tiles=(64, 64), receivers=256, spacing=10.0, source_median=500, building_median=300,
                     sigma=1.0, source_clip=(10, 50_000), building_clip=(1, 50_000)):
    """Configuration "batched uneven tiles" at its full size, generated vectorized: as
    tile_mosaic (log-normal per-tile source and building counts, one diamond-square DEM at 30 m)
    with sources at uniform places in their tile and receivers standing outdoors everywhere (no
    receiver takes a building as its own footprint)."""
    rng = np.random.default_rng(seed)
    tx, ty = tiles
    side = receivers * spacing
    extent = (tx * side, ty * side)
    margin = 600.0
    nodes_needed = int(math.ceil((max(extent) + 2 * margin) / 30.0)) + 1
    power = max(int(math.ceil(math.log2(max(nodes_needed - 1, 2)))), 2)
    dem = _diamond_square(rng, power, 200.0)[:nodes_needed, :nodes_needed] + 100.0
    impervious = np.clip(rng.normal(40, 30, dem.shape), 0, 100).astype(np.uint32)
    terrain = abi.Terrain(-margin, -margin, 30.0, abi.Terrain.pack(dem, impervious))
    del dem, impervious
    n_tiles = tx * ty
    n_src = np.clip(np.round(source_median * np.exp(sigma * rng.normal(size=n_tiles))), *source_clip).astype(np.int64)
    n_bld = np.clip(np.round(building_median * np.exp(sigma * rng.normal(size=n_tiles))), *building_clip).astype(np.int64)
    tile_x, tile_y = (np.arange(n_tiles) % tx) * side, (np.arange(n_tiles) // tx) * side
    # Buildings: 4-12-gons of radius 5-20 m about uniform centres in their tile.
    owner = np.repeat(np.arange(n_tiles), n_bld)
    nb = owner.size
    centre = rng.uniform(0, side, (nb, 2)) + np.stack([tile_x[owner], tile_y[owner]], 1)
    k = rng.integers(4, 13, nb)
    b = np.repeat(np.arange(nb), k)
    j = np.arange(b.size) - np.concatenate([[0], np.cumsum(k)[:-1]])[b]
    # Sorted angles per polygon: evenly spaced plus jitter within each sector.
    ang = (j + rng.uniform(0.1, 0.9, b.size)) * (2 * math.pi / k[b])
    rad = rng.uniform(5.0, 20.0, nb)[b] * rng.uniform(0.6, 1.0, b.size)
    xy = np.stack([centre[b, 0] + rad * np.cos(ang), centre[b, 1] + rad * np.sin(ang)], 1)
    obstacles = _obstacles_fast(xy, k, rng.uniform(3.0, 60.0, nb), (0.0, 0.0), extent)
    del xy, b, j, ang, rad
    weather = _weather((-margin, -margin), (extent[0] + 2 * margin, extent[1] + 2 * margin),
                       lambda r, c: np.full((3, 16), 50))
    owner_s = np.repeat(np.arange(n_tiles), n_src)
    src = np.zeros(owner_s.size, abi.SOURCE_DTYPE)
    src["x_m"] = rng.uniform(0, side, owner_s.size) + tile_x[owner_s]
    src["y_m"] = rng.uniform(0, side, owner_s.size) + tile_y[owner_s]
    rail = rng.uniform(0, 1, owner_s.size) < 0.3
    src["height_m"] = np.where(rail, 0.5, 0.05)
    src["layer"] = np.where(rail, 1, 0)
    src["exclusion_radius_m"] = 2.0
    src["ground_percent"] = abi.GROUND_FROM_TERRAIN
    levels = rng.normal(85.0, 8.0, (owner_s.size, 8)) + A_WEIGHTING_DB
    src["energy"] = 10.0 ** ((levels[:, None, :] + np.array([0.0, -2.0, -7.0])[None, :, None]) / 10.0)
    first = np.concatenate([[0], np.cumsum(n_src)]).astype(np.uint32)
    # Receivers tile by tile (each tile's 256 x 256 block contiguous).
    ii, jj = np.meshgrid(np.arange(receivers), np.arange(receivers))
    local = np.stack([(ii.ravel() + 0.5) * spacing, (jj.ravel() + 0.5) * spacing], 1)
    pts = np.zeros(n_tiles * receivers * receivers, abi.POINT_DTYPE)
    per = receivers * receivers
    for t in range(n_tiles):
        pts["x_m"][t * per:(t + 1) * per] = local[:, 0] + tile_x[t]
        pts["y_m"][t * per:(t + 1) * per] = local[:, 1] + tile_y[t]
    scene = abi.Scene(terrain, obstacles, weather)
    return TileMosaic(scene, src, pts, (receivers, receivers), (tx, ty), first)


def line_pieces(rng, count_lines, extent, origin, piece_m=250.0, height_m=0.05, level_mean=75.0, level_sd=6.0):
    """Random polylines (2-5 vertices over the extent) cut into straight pieces of at most
    `piece_m`, with an emission per metre of N(level_mean, level_sd) dB per band + A-weighting."""
    out = []
    for _ in range(count_lines):
        k = int(rng.integers(2, 6))
        verts = np.stack([origin[0] + rng.uniform(0, extent[0], k), origin[1] + rng.uniform(0, extent[1], k)], 1)
        levels = rng.normal(level_mean, level_sd, 8) + A_WEIGHTING_DB
        for a, b in zip(verts[:-1], verts[1:]):
            length = float(np.hypot(*(b - a)))
            parts = max(int(math.ceil(length / piece_m)), 1)
            for j in range(parts):
                p = np.zeros((), abi.LINE_PIECE_DTYPE)
                p["ax_m"], p["ay_m"] = a + (b - a) * (j / parts)
                p["bx_m"], p["by_m"] = a + (b - a) * ((j + 1) / parts)
                p["height_m"] = height_m
                p["energy_per_m"] = np.repeat((10.0 ** (levels / 10.0))[None, :], 3, axis=0)
                out.append(p)
    return np.array(out, abi.LINE_PIECE_DTYPE)


def line_scene(seed=5, receivers=96, spacing=10.0, lines=30, buildings=300):
    """The plain road scene's ground, buildings and weather with road polylines kept as line
    pieces (<= 250 m) instead of points: the line-source rows' test scene."""
    sc = road_scene(seed=seed, receivers=receivers, spacing=spacing, lines=1, per_line=1, buildings=buildings)
    rng = np.random.default_rng(seed + 100)
    extent = receivers * spacing
    sc.pieces = line_pieces(rng, lines, (extent, extent), (0.0, 0.0))
    return sc


TERRAIN_BAND_LIMITS_M = (500.0, 1000.0, 2000.0, 3500.0, 5500.0, 8000.0)
BUILDING_BAND_LIMITS_M = (16.0, 32.0, 64.0, 128.0, 256.0, 512.0)


def _sector_edges(rng, sectors, limits, tangent_lo, tangent_hi, empty):
    """One SectorEdges as the card reads it: steepest of all, then per sector its steepest and its
    bands' (range, tangent), NaN range for an empty band."""
    out = [0.0]
    steepest_all = -np.inf
    lower = (0.0,) + limits[:-1]
    for _ in range(sectors):
        bands, steepest = [], -np.inf
        for lo, hi in zip(lower, limits):
            if rng.uniform() < empty:
                bands += [np.nan, np.nan]
                continue
            r = rng.uniform(max(lo, 1.0), hi)
            t = rng.uniform(tangent_lo, tangent_hi)
            bands += [r, t]
            steepest = max(steepest, t)
        out += [steepest] + bands
        steepest_all = max(steepest_all, steepest)
    out[0] = steepest_all
    return out


def aircraft_horizons(scene, seed=9, block=8):
    """Synthetic horizons for the aircraft grid: one record a `block` x `block` receivers (terrain
    ridges to a tangent of 0.06 within 8 km, buildings to a tangent of 1.2 within 512 m, a tenth of
    the blocks unscreened), and the terrain under every vertex (0 to 150 m). A dict for
    qm_aircraft_screening."""
    rng = np.random.default_rng(seed)
    nx, ny = scene.grid_shape
    bx, by = (nx + block - 1) // block, (ny + block - 1) // block
    records = []
    for _ in range(bx * by):
        rec = _sector_edges(rng, abi.TERRAIN_SECTORS, TERRAIN_BAND_LIMITS_M, -0.02, 0.06, 0.2)
        rec += _sector_edges(rng, abi.BUILDING_SECTORS, BUILDING_BAND_LIMITS_M, 0.05, 1.2, 0.6)
        records.append(rec)
    horizons = np.array(records, np.float32)
    assert horizons.shape[1] == abi.HORIZON_FLOATS
    jj, ii = np.mgrid[0:ny, 0:nx]
    of = ((jj // block) * bx + ii // block).reshape(-1).astype(np.uint32)
    of[rng.uniform(size=of.size) < 0.1] = abi.NO_HORIZON
    ground = rng.uniform(0.0, 150.0, scene.vertices.reshape(-1, 3).shape[0]).astype(np.float32)
    return dict(horizons=horizons, horizon_of=of, vertex_ground_m=ground)


def with_fine_terrain(terrain, seed=13, share=0.5, amplitude_m=4.0):
    """`terrain` (abi.Terrain) with fine terrain: seeded detail of up to `amplitude_m` (smooth
    ripples plus noise, in QM_DETAIL_STEP_M steps) on the lattice six times as fine, read in
    `share` of its 32-cell blocks."""
    rng = np.random.default_rng(seed)
    rows, columns = terrain.nodes.shape
    fr, fc = abi.FINE_PER_NODE * (rows - 1) + 1, abi.FINE_PER_NODE * (columns - 1) + 1
    yy, xx = np.mgrid[0:fr, 0:fc].astype(np.float64)
    ripple = np.sin(xx / 7.3 + rng.uniform(0, 6)) * np.cos(yy / 5.1 + rng.uniform(0, 6))
    detail_m = amplitude_m * (0.7 * ripple + 0.3 * rng.uniform(-1, 1, (fr, fc)))
    detail = np.round(detail_m / abi.DETAIL_STEP_M).astype(np.int16)
    blocks = (rng.uniform(size=((rows - 2) // abi.FINE_BLOCK_CELLS + 1, (columns - 2) // abi.FINE_BLOCK_CELLS + 1))
              < share).astype(np.uint8)
    return abi.Terrain(terrain.origin_x_m, terrain.origin_y_m, terrain.spacing_m, terrain.nodes, terrain.canopy_m,
                       fine_detail=detail, fine_blocks=blocks, column_m=terrain.column_m, row_y_m=terrain.row_y_m)
