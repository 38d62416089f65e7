"""Aircraft boxes as the card reads them: host records (qm_aircraft.h qm_flight_box) of decoded
boxes, as the reference's host writes them for its card (paint-gpu/src/aircraft.rs:160-205
DeviceFlightBoxes::upload: the centroid in the square's frame, the axis as its sine and cosine,
the average piece's length, the passing shares, the FP32 levels, lg d_lambda and installation
shares, the screened, helicopter and silence flags), and seeded synthetic boxes for the parity
tests and the profile workload.
"""
from __future__ import annotations

import math

import numpy as np

from . import abi

NPD_DISTANCES_FT = np.array([200.0, 400.0, 630.0, 1000.0, 2000.0, 4000.0, 6300.0, 10000.0, 16000.0, 25000.0])
METRES_PER_FOOT = 0.3048
TAIL_ANCHOR_M = 16000.0
GEOMETRY_DISTANCE = 3  # 1,000 ft
SCREENING_CEILING_M = 7620.0
PERIOD_SECONDS = np.array([12.0, 4.0, 8.0]) * 3600.0


def average_length_m(levels_db, piece_length_m):
    """physics/src/doc29/boxes.rs:99-111: the periods' piece lengths weighted by their energy at
    1,000 ft."""
    levels = np.asarray(levels_db, np.float64)
    w = np.where(np.isfinite(levels[:, GEOMETRY_DISTANCE]), 10.0 ** (levels[:, GEOMETRY_DISTANCE] / 10.0), 0.0)
    total = w.sum()
    return float((w * np.asarray(piece_length_m)).sum() / total) if total > 0 else float(piece_length_m[0])


def box_record(values, passing, ground_m=0.0, helicopter=False, east_m=None, north_m=None, altitude_m=None):
    """A box's record from its values (levels_db [3][10], tail_levels_db [3], scaled_distance_m
    [10], centroid_m, axis_rad, gradient, gradient_spread, piece_length_m [3],
    installation_shares [3][2][3]) in the square's frame, its passing shares per period and the
    ground under it."""
    r = np.zeros(1, abi.FLIGHT_BOX_DTYPE)
    c = values["centroid_m"]
    r["east_m"] = c[0] if east_m is None else east_m
    r["north_m"] = c[1] if north_m is None else north_m
    r["centroid_altitude_m"] = c[2] if altitude_m is None else altitude_m
    r["sin_axis"], r["cos_axis"] = math.sin(values["axis_rad"]), math.cos(values["axis_rad"])
    r["gradient"], r["gradient_spread"] = values["gradient"], values["gradient_spread"]
    r["average_length_m"] = average_length_m(values["levels_db"], values["piece_length_m"])
    r["passing"] = passing
    r["piece_length_m"] = values["piece_length_m"]
    levels = np.asarray(values["levels_db"], np.float64)
    r["levels_db"] = levels
    r["tail_levels_db"] = values["tail_levels_db"]
    r["lg_scaled_distance"] = np.log10(values["scaled_distance_m"])
    r["installation_shares"] = values["installation_shares"]
    r["screened"] = int(r["centroid_altitude_m"][0] - ground_m < SCREENING_CEILING_M)
    r["helicopter"] = int(bool(helicopter))
    r["silent"] = np.all(levels == -np.inf, axis=1).astype(np.uint32)
    return r


def values_of_record(r):
    """The box values a record holds (its FP32 fields as they are), the centroid in the frame."""
    return dict(
        levels_db=np.where(r["silent"][:, None] != 0, -np.inf, r["levels_db"].astype(np.float64)),
        tail_levels_db=r["tail_levels_db"].astype(np.float64),
        scaled_distance_m=10.0 ** r["lg_scaled_distance"].astype(np.float64),
        centroid_m=np.array([r["east_m"], r["north_m"], r["centroid_altitude_m"]], np.float64),
        axis_rad=math.atan2(float(r["sin_axis"]), float(r["cos_axis"])),
        gradient=float(r["gradient"]),
        gradient_spread=float(r["gradient_spread"]),
        piece_length_m=r["piece_length_m"].astype(np.float64),
        installation_shares=r["installation_shares"].astype(np.float64),
        average_length_m=float(r["average_length_m"]),
    )


def synthetic_boxes(seed=11, count=4096, extent_m=20000.0, origin=(0.0, 0.0), silent_share=0.1):
    """Seeded boxes over a square of `extent_m` from `origin`: centroids 150 m to 3 km up, axes and
    gradients of climbs and descents (a third with a spread of gradients), NPD-like summed levels
    (SEL at 1,000 ft ~ N(88, 8) dB, 13-19 dB a decade, 0.6-1.2 dB a kft of absorption) per period
    with some silent periods, lg d_lambda of a 150-180 kt reference speed, mixed installation
    shares, passing shares of 1e-5 to 1e-2, one box in 20 a helicopter's."""
    rng = np.random.default_rng(seed)
    r = np.zeros(count, abi.FLIGHT_BOX_DTYPE)
    r["east_m"] = origin[0] + rng.uniform(0.0, extent_m, count)
    r["north_m"] = origin[1] + rng.uniform(0.0, extent_m, count)
    r["centroid_altitude_m"] = rng.uniform(150.0, 3000.0, count)
    axis = rng.uniform(0.0, math.pi, count)
    r["sin_axis"], r["cos_axis"] = np.sin(axis), np.cos(axis)
    r["gradient"] = rng.normal(0.0, 0.06, count)
    r["gradient_spread"] = np.where(rng.uniform(size=count) < 0.33, rng.uniform(0.01, 0.06, count), 0.0)
    lengths = rng.uniform(80.0, 400.0, (count, 3))
    r["piece_length_m"] = lengths
    d_ft = NPD_DISTANCES_FT
    base = rng.normal(88.0, 8.0, count)
    slope = rng.uniform(13.0, 19.0, count)
    absorb = rng.uniform(0.0006, 0.0012, count)
    curve = base[:, None] - slope[:, None] * np.log10(d_ft / 1000.0) - absorb[:, None] * (d_ft - 1000.0)
    tail_ft = TAIL_ANCHOR_M / METRES_PER_FOOT
    tail = base - slope * math.log10(tail_ft / 1000.0) - absorb * (tail_ft - 1000.0)
    offsets = np.stack([np.zeros(count), rng.uniform(-8.0, 0.0, count), rng.uniform(-12.0, 0.0, count)], 1)
    silent = rng.uniform(size=(count, 3)) < silent_share
    silent[:, 0] = False
    levels = curve[:, None, :] + offsets[:, :, None]
    levels[silent] = -np.inf
    r["levels_db"] = levels
    r["tail_levels_db"] = np.where(silent, -np.inf, tail[:, None] + offsets)
    r["silent"] = silent.astype(np.uint32)
    v_ref = rng.uniform(140.0, 180.0, count) * 1852.0 / 3600.0
    delta = 8.0 + 6.0 * np.log10(d_ft / 1000.0)[None, :] + rng.uniform(-1.0, 1.0, (count, 1))
    r["lg_scaled_distance"] = np.log10(2.0 / math.pi * v_ref)[:, None] + delta / 10.0
    shares = rng.dirichlet([1.0, 0.5, 0.5], (count, 3, 2))
    r["installation_shares"] = shares
    r["passing"] = 10.0 ** rng.uniform(-5.0, -2.0, (count, 3))
    r["passing"][silent] = 0.0
    r["screened"] = 1
    r["helicopter"] = (rng.uniform(size=count) < 0.05).astype(np.uint32)
    for i in range(count):
        r["average_length_m"][i] = average_length_m(levels[i], lengths[i])
    return r


def flight_requests(boxes, points_xy, altitude_m=4.0, tiles=(4, 4), extent_m=20000.0, origin=(0.0, 0.0),
                    within_m=16000.0, weights=(abi.FLIGHT_WEIGHT_ONE,), coarse_m=500.0, horizons=False):
    """Points and segments over boxes sorted into `tiles` square tiles: each point hears every
    tile (a segment a tile and weight, into sum slot point * len(weights) + k). Returns (sorted
    boxes, points, segments, n_sums)."""
    tx = np.clip(((boxes["east_m"] - origin[0]) / extent_m * tiles[0]).astype(int), 0, tiles[0] - 1)
    ty = np.clip(((boxes["north_m"] - origin[1]) / extent_m * tiles[1]).astype(int), 0, tiles[1] - 1)
    tile = ty * tiles[0] + tx
    order = np.argsort(tile, kind="stable")
    boxes = boxes[order]
    tile = tile[order]
    first = np.searchsorted(tile, np.arange(tiles[0] * tiles[1]))
    counts = np.bincount(tile, minlength=tiles[0] * tiles[1])
    xy = np.asarray(points_xy, np.float64).reshape(-1, 2)
    points = np.zeros(xy.shape[0], abi.FLIGHT_POINT_DTYPE)
    points["x"], points["y"] = xy[:, 0], xy[:, 1]
    points["altitude_m"] = altitude_m
    points["horizons"] = np.arange(xy.shape[0], dtype=np.uint64) * abi.HORIZON_FLOATS if horizons else abi.NO_HORIZONS
    points["impedance_gain"] = 1.0
    segs = []
    for p in range(xy.shape[0]):
        for k, w in enumerate(weights):
            for t in range(tiles[0] * tiles[1]):
                if counts[t]:
                    segs.append((within_m, p, p * len(weights) + k, first[t], counts[t], w, coarse_m))
    segments = np.array(segs, abi.FLIGHT_SEGMENT_DTYPE)
    return boxes, points, segments, xy.shape[0] * len(weights)
