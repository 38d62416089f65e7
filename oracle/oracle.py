"""TEST INFRASTRUCTURE — ctypes binding of the FP64 CPU oracle (``liboracle.so``).

Only the tests, ``__graft_entry__.smoke()`` and the bench's CPU-baseline leg import this module,
and only to check or time against it; the product path (``quietmap_amd``) never does.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

from quietmap_amd import abi

HERE = os.path.dirname(os.path.abspath(__file__))
LIBRARY = os.path.join(HERE, "liboracle.so")

_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE, "liboracle.so"], check=True)


NATIVE = os.path.join(HERE, "_native", "liboracle.so")
NATIVE_FLAGS = "g++ -O3 -march=native -ffp-contract=off"


def use_native():
    """Load the CPU baseline's build (NATIVE_FLAGS, compiled on this host for its own CPU) in place of
    the checker build; both compute the same (no fast math, no contraction)."""
    global LIBRARY, _lib
    if LIBRARY == NATIVE and _lib is not None:
        return
    subprocess.run(["make", "-s", "-j8", "-C", HERE, "_native/liboracle.so"], check=True)
    LIBRARY, _lib = NATIVE, None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIBRARY):
            build()
        _lib = C.CDLL(LIBRARY)
        _declare(_lib)
    return _lib


D = C.POINTER(C.c_double)
U8 = C.POINTER(C.c_uint8)
U32 = C.POINTER(C.c_uint32)
U64 = C.POINTER(C.c_uint64)
F32 = C.POINTER(C.c_float)
SZ = C.POINTER(C.c_size_t)


def _declare(L):
    L.qmo_alpha_bands.argtypes = [C.c_double, C.c_double, D]
    L.qmo_lden_energy.argtypes = [D]
    L.qmo_lden_energy.restype = C.c_double
    L.qmo_state_boundaries.argtypes = [C.c_size_t, D, D, D, C.c_double, C.c_double, C.c_double, C.c_double,
                                       C.c_double, C.c_size_t, D, D, C.c_size_t, D, D, D, D]
    L.qmo_ray_terms.argtypes = [C.c_size_t, C.c_double, D, D, D, D, D, C.c_size_t, D, D, U8, U8, U64,
                                C.c_double, C.c_int, C.c_double, C.c_double, C.c_int, C.c_double, C.c_double,
                                C.c_uint64, D, D, D, D, D, D]
    L.qmo_vertical_path.argtypes = [C.c_size_t, C.c_double, D, D, D, C.c_size_t, D, D, U8, U8, U64, C.c_double,
                                    C.c_double, C.c_uint64, C.c_int, C.c_double, C.c_size_t, D, D, D, SZ, D, SZ, D,
                                    SZ]
    L.qmo_ground_at.argtypes = [C.c_void_p, C.c_double, C.c_double, D]
    L.qmo_fill_profile.argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_double, C.c_double, C.c_size_t,
                                   D, D, D, D, D, SZ, D]
    L.qmo_crossings.argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_double, C.c_double, C.c_size_t,
                                D, D, U8, U8, U64, U32, U32, SZ, C.c_int]
    L.qmo_receivers.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t, D, D, D, U32]
    L.qmo_evaluate_pairs.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t, C.c_void_p,
                                     C.c_size_t, U32, U32, F32, C.c_size_t, D, U32, D, C.c_int, U64]
    L.qmo_line_nodes.argtypes = [D, D, D, C.c_size_t, C.c_size_t, D, D, U8, D]
    L.qmo_line_nodes.restype = C.c_long
    L.qmo_skyline_arcs.argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_double, C.c_double, C.c_double,
                                   C.c_double, C.c_size_t, D]
    L.qmo_skyline_arcs.restype = C.c_long
    L.qmo_evaluate_line_pairs.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t, C.c_void_p,
                                          C.c_size_t, U32, U32, C.c_size_t, D, U32, U32, C.c_int]
    L.qmo_pair_boundaries.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t, C.c_void_p,
                                      C.c_size_t, U32, U32, C.c_size_t, D, U32, C.c_int]
    L.qmo_paint_points.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t, C.c_void_p,
                                   C.c_size_t, C.c_double, D, U64, U32, C.c_int]
    L.qmo_zwicker.argtypes = [D, C.c_size_t, D]
    L.qmo_kurze_table.argtypes = [F32]
    L.qmo_relative_intensity.argtypes = [C.c_double, C.c_double]
    L.qmo_relative_intensity.restype = C.c_double
    L.qmo_node_probabilities.argtypes = [D, C.c_size_t]
    L.qmo_node_probabilities.restype = C.c_size_t
    L.qmo_line_values.argtypes = [C.c_int, C.c_double, D, C.c_size_t]
    L.qmo_line_values.restype = C.c_size_t
    L.qmo_runs.argtypes = [D, D, C.c_size_t, C.c_double, C.c_double, C.c_double, D]
    L.qmo_runs.restype = C.c_int
    L.qmo_many_lines_energy.argtypes = [C.c_double, C.c_int, C.c_double]
    L.qmo_many_lines_energy.restype = C.c_double
    L.qmo_line_exceeded_db.argtypes = [C.c_double, C.c_double, C.c_double]
    L.qmo_line_exceeded_db.restype = C.c_double
    L.qmo_hour_groups.argtypes = [C.POINTER(C.c_int), D]
    L.qmo_pixel_loudness.argtypes = [F32, U64, C.c_size_t, C.c_double, D, D, C.c_int]
    L.qmo_fold_pixels.argtypes = [D, U64, C.c_size_t, C.c_double, F32, C.c_void_p, U64]
    L.qmo_fold_pixels.restype = None
    L.qmo_merge_close_states.argtypes = [D, C.POINTER(C.c_uint32), C.c_size_t, C.c_double]
    L.qmo_merge_close_states.restype = None
    L.qmo_finite_fraction.argtypes = [C.c_double, C.c_double, C.c_double]
    L.qmo_finite_fraction.restype = C.c_double
    L.qmo_lateral_attenuation.argtypes = [C.c_double, C.c_double]
    L.qmo_lateral_attenuation.restype = C.c_double
    L.qmo_installation_gain.argtypes = [C.c_uint32, C.c_double, C.c_double]
    L.qmo_installation_gain.restype = C.c_double
    L.qmo_closest_points.argtypes = [D, D, D]
    L.qmo_segment_levels.argtypes = [C.c_void_p, C.c_void_p, D, D, D, D]
    L.qmo_npd_distance_levels.argtypes = [C.c_void_p, C.c_void_p, D]
    L.qmo_box_values.argtypes = [C.c_void_p, C.c_void_p, D, D, C.c_size_t, D]
    L.qmo_box_sel.argtypes = [D, D, F32, D, D]
    L.qmo_flight_sums.argtypes = [C.c_void_p, C.c_size_t, C.c_void_p, C.c_size_t, F32, C.c_size_t, C.c_void_p,
                                  C.c_size_t, D, C.c_size_t]
    L.qmo_aircraft_points.argtypes = [C.c_void_p, C.c_size_t, C.c_void_p, C.c_size_t, C.c_void_p, D, C.c_size_t,
                                      C.c_double, C.c_double, D, U32, U32, U32, D, C.c_void_p, U64, C.c_int]
    L.qmo_path_difference_loss.argtypes = [C.c_double]
    L.qmo_path_difference_loss.restype = C.c_double
    L.qmo_edge_loss.argtypes = [C.c_double] * 4
    L.qmo_edge_loss.restype = C.c_double
    L.qmo_sector_of.argtypes = [C.c_double, C.c_double, C.c_int]
    L.qmo_sector_of.restype = C.c_int
    L.qmo_horizon_loss.argtypes = [F32, C.c_int, C.c_double, C.c_double, C.c_double]
    L.qmo_horizon_loss.restype = C.c_double
    L.qmo_horizons.argtypes = [C.c_void_p, C.c_void_p, C.c_double, C.c_double, C.c_double, C.c_uint64, F32]
    L.qmo_horizons.restype = C.c_int


def finite_fraction(along, length, scaled):
    return lib().qmo_finite_fraction(along, length, scaled)


def lateral_attenuation(height, lateral):
    return lib().qmo_lateral_attenuation(height, lateral)


def installation_gain(installation, height, slant):
    return lib().qmo_installation_gain(installation, height, slant)


def closest_points(start, end):
    s = np.ascontiguousarray(start, np.float64)
    e = np.ascontiguousarray(end, np.float64)
    out = np.zeros(12)
    lib().qmo_closest_points(_p(s, D), _p(e, D), _p(out, D))
    return dict(along=out[0], on_line=out[1:4], on_segment=out[4:7], lateral_m=out[7], height_m=out[8],
                lambda_height_m=out[9], length_m=out[10])


def segment_levels(classes, track, start, end):
    """SEL and LAmax (dB) of one segment of `track` in the receiver's frame."""
    classes = np.ascontiguousarray(classes, abi.AIRCRAFT_CLASS_DTYPE)
    track = np.ascontiguousarray(np.asarray(track, abi.TRACK_DTYPE).reshape(1))
    s = np.ascontiguousarray(start, np.float64)
    e = np.ascontiguousarray(end, np.float64)
    sel, lmax = np.zeros(1), np.zeros(1)
    lib().qmo_segment_levels(classes.ctypes.data, track.ctypes.data, _p(s, D), _p(e, D), _p(sel, D), _p(lmax, D))
    return sel[0], lmax[0]


class Screening(C.Structure):
    _fields_ = [("horizons", C.c_void_p), ("n_horizons", C.c_size_t), ("horizon_of", C.c_void_p),
                ("vertex_ground_m", C.c_void_p)]


def screening_struct(screening):
    """A qm_aircraft_screening over a dict of `horizons` (n, QM_HORIZON_FLOATS) float32,
    `horizon_of` (per receiver, uint32) and optional `vertex_ground_m` (float32); the arrays it
    points at are returned too, to be kept alive."""
    h = np.ascontiguousarray(screening["horizons"], np.float32)
    of = np.ascontiguousarray(screening["horizon_of"], np.uint32)
    g = screening.get("vertex_ground_m")
    g = None if g is None else np.ascontiguousarray(g, np.float32)
    st = Screening(h.ctypes.data, h.shape[0], of.ctypes.data, None if g is None else g.ctypes.data)
    return st, (h, of, g)


def aircraft_points(scene, points_xy, threads=None, passes=False, screening=None):
    """Per receiver: SEL energy per period, N65, N70, LAmax histogram, segments heard, and with
    `passes` the pass classes [n, 2, 3, 48] (energies, then passing shares). `screening`: a dict
    (`screening_struct`) whose `horizon_of` is indexed by these points."""
    xy = np.ascontiguousarray(points_xy, np.float64).reshape(-1, 2)
    n = xy.shape[0]
    energy = np.zeros((n, 3))
    n65, n70 = np.zeros(n, np.uint32), np.zeros(n, np.uint32)
    hist = np.zeros((n, abi.HIST_BINS), np.uint32)
    heard = np.zeros(n, np.uint64)
    pc = np.zeros((n, 2, 3, abi.PASS_CLASSES)) if passes else None
    if screening:
        st, keep = screening_struct(screening)
    classes = np.ascontiguousarray(scene.classes, abi.AIRCRAFT_CLASS_DTYPE)
    tracks = np.ascontiguousarray(scene.tracks, abi.TRACK_DTYPE)
    verts = np.ascontiguousarray(scene.vertices, np.float32)
    rc = lib().qmo_aircraft_points(classes.ctypes.data, classes.size, tracks.ctypes.data, tracks.size,
                                   verts.ctypes.data, _p(xy, D), n, scene.receiver_altitude_m, scene.reach_m,
                                   _p(energy, D), _p(n65, U32), _p(n70, U32), _p(hist, U32),
                                   _p(pc, D) if passes else None, C.byref(st) if screening else None,
                                   _p(heard, U64), threads or os.cpu_count() or 1)
    if rc != 0:
        raise ValueError("oracle: track out of its class's power rows")
    out = dict(period_energy=energy, n65=n65, n70=n70, hist=hist, segments=heard)
    if passes:
        out["passes"] = pc
    return out


def zwicker_sone(levels):
    """ISO 532-1 stationary loudness (sone) of third-octave levels (..., 28) in dB."""
    lv = np.ascontiguousarray(np.asarray(levels, np.float64).reshape(-1, 28))
    out = np.zeros(lv.shape[0])
    lib().qmo_zwicker(_p(lv, D), lv.shape[0], _p(out, D))
    return out


def _p(a, t):
    return a.ctypes.data_as(t)


def alpha_bands(temperature_c, relative_humidity_pct):
    out = np.zeros(8)
    lib().qmo_alpha_bands(temperature_c, relative_humidity_pct, _p(out, D))
    return out


def lden_energy(period_energy):
    e = np.ascontiguousarray(period_energy, np.float64)
    return lib().qmo_lden_energy(_p(e, D))


def state_boundaries(distance, altitude, ground, source, receiver, source_ground_factor, terrain, tops):
    d, z, g = (np.ascontiguousarray(a, np.float64) for a in (distance, altitude, ground))
    terrain = np.ascontiguousarray(np.asarray(terrain, np.float64).reshape(-1, 2))
    tops = np.ascontiguousarray(np.asarray(tops, np.float64).reshape(-1, 2))
    tx, tz = np.ascontiguousarray(terrain[:, 0]), np.ascontiguousarray(terrain[:, 1])
    ox, oz = np.ascontiguousarray(tops[:, 0]), np.ascontiguousarray(tops[:, 1])
    att, wo = np.zeros(16), np.zeros(16)
    lib().qmo_state_boundaries(d.size, _p(d, D), _p(z, D), _p(g, D), source[0], source[1], receiver[0], receiver[1],
                               source_ground_factor, tx.size, _p(tx, D), _p(tz, D), ox.size, _p(ox, D), _p(oz, D),
                               _p(att, D), _p(wo, D))
    return att.reshape(2, 8), wo.reshape(2, 8)


def ray_terms(profile, crossings, ends, favourable, alpha):
    """profile: dict(horizontal_m, t, ground_m, ground_factor, forest_cover, canopy_m);
    crossings: list of (t, height_m, building, bounds_roof, footprint_id); ends: dict."""
    arr = {k: np.ascontiguousarray(profile[k], np.float64) for k in ("t", "ground_m", "ground_factor")}
    n = arr["t"].size
    forest = np.ascontiguousarray(profile.get("forest_cover", np.zeros(n)), np.float64)
    canopy = np.ascontiguousarray(profile.get("canopy_m", np.zeros(n)), np.float64)
    ct = np.array([c[0] for c in crossings], np.float64)
    ch = np.array([c[1] for c in crossings], np.float64)
    cb = np.array([bool(c[2]) for c in crossings], np.uint8)
    cr = np.array([bool(c[3]) for c in crossings], np.uint8)
    cf = np.array([c[4] for c in crossings], np.uint64)
    fav = np.ascontiguousarray(favourable, np.float64)
    al = np.ascontiguousarray(alpha, np.float64)
    periods, states, slant, boundary = np.zeros(24), np.zeros(16), np.zeros(1), np.zeros(16)
    sg = ends.get("source_ground_m")
    sgf = ends.get("source_ground_factor")
    lib().qmo_ray_terms(
        n, profile["horizontal_m"], _p(arr["t"], D), _p(arr["ground_m"], D), _p(arr["ground_factor"], D),
        _p(forest, D), _p(canopy, D), ct.size, _p(ct, D), _p(ch, D), _p(cb, U8), _p(cr, U8), _p(cf, U64),
        ends["source_height_m"], int(sg is not None), 0.0 if sg is None else sg, ends["receiver_altitude_m"],
        int(sgf is not None), 0.0 if sgf is None else sgf, ends.get("platform_half_width_m", 0.0),
        ends.get("own_footprint", 0), _p(fav, D), _p(al, D), _p(periods, D), _p(states, D), _p(slant, D),
        _p(boundary, D))
    return dict(periods=periods.reshape(3, 8), states=states.reshape(2, 8), slant_m=slant[0],
                boundary_db=boundary.reshape(2, 8))


def vertical_path(profile, crossings, source_height_m, platform_half_width_m=0.0, own_footprint=0, capacity=4096,
                  source_ground_m=None):
    arr = {k: np.ascontiguousarray(profile[k], np.float64) for k in ("t", "ground_m", "ground_factor")}
    ct = np.array([c[0] for c in crossings], np.float64)
    ch = np.array([c[1] for c in crossings], np.float64)
    cb = np.array([bool(c[2]) for c in crossings], np.uint8)
    cr = np.array([bool(c[3]) for c in crossings], np.uint8)
    cf = np.array([c[4] for c in crossings], np.uint64)
    d, z, g = np.zeros(capacity), np.zeros(capacity), np.zeros(capacity)
    tops, crests = np.zeros(2 * capacity), np.zeros(2 * capacity)
    n, nt, nc = C.c_size_t(0), C.c_size_t(0), C.c_size_t(0)
    rc = lib().qmo_vertical_path(arr["t"].size, profile["horizontal_m"], _p(arr["t"], D), _p(arr["ground_m"], D),
                                 _p(arr["ground_factor"], D), ct.size, _p(ct, D), _p(ch, D), _p(cb, U8), _p(cr, U8),
                                 _p(cf, U64), source_height_m, platform_half_width_m, own_footprint,
                                 int(source_ground_m is not None), 0.0 if source_ground_m is None else source_ground_m,
                                 capacity,
                                 _p(d, D), _p(z, D), _p(g, D), C.byref(n), _p(tops, D), C.byref(nt), _p(crests, D),
                                 C.byref(nc))
    assert rc == 0
    k = n.value
    return dict(d=d[:k], z=z[:k], g=g[:k], tops=tops[: 2 * nt.value].reshape(-1, 2),
                crests=crests[: 2 * nc.value].reshape(-1, 2))


def ground_at(terrain: abi.Terrain, x, y):
    out = np.zeros(4)
    failed = lib().qmo_ground_at(C.byref(terrain.c), x, y, _p(out, D))
    return failed, out


def fill_profile(terrain: abi.Terrain, source, receiver, capacity=1 << 16):
    bufs = [np.zeros(capacity) for _ in range(5)]
    n, horizontal = C.c_size_t(0), np.zeros(1)
    failed = lib().qmo_fill_profile(C.byref(terrain.c), source[0], source[1], receiver[0], receiver[1], capacity,
                                    *[_p(b, D) for b in bufs], C.byref(n), _p(horizontal, D))
    k = n.value
    return failed, dict(horizontal_m=horizontal[0], t=bufs[0][:k], ground_m=bufs[1][:k], ground_factor=bufs[2][:k],
                        forest_cover=bufs[3][:k], canopy_m=bufs[4][:k])


CROSSING_DTYPE = np.dtype([("t", "<f8"), ("height_m", "<f8"), ("building", "u1"), ("bounds_roof", "u1"),
                           ("footprint_id", "<u8"), ("outline", "<u4"), ("edge", "<u4")])


def crossings(obstacles: abi.Obstacles, start, end, brute_force=False, capacity=4096):
    t, h = np.zeros(capacity), np.zeros(capacity)
    b, r = np.zeros(capacity, np.uint8), np.zeros(capacity, np.uint8)
    f = np.zeros(capacity, np.uint64)
    o, e = np.zeros(capacity, np.uint32), np.zeros(capacity, np.uint32)
    n = C.c_size_t(0)
    failed = lib().qmo_crossings(C.byref(obstacles.c), start[0], start[1], end[0], end[1], capacity, _p(t, D),
                                 _p(h, D), _p(b, U8), _p(r, U8), _p(f, U64), _p(o, U32), _p(e, U32), C.byref(n),
                                 int(brute_force))
    k = n.value
    out = np.zeros(k, CROSSING_DTYPE)
    if failed == 0:
        out["t"], out["height_m"], out["building"], out["bounds_roof"] = t[:k], h[:k], b[:k], r[:k]
        out["footprint_id"], out["outline"], out["edge"] = f[:k], o[:k], e[:k]
    return failed, out


def receivers(scene: abi.Scene, points):
    points = np.ascontiguousarray(points, abi.POINT_DTYPE)
    n = points.size
    altitude, alpha, fav = np.zeros(n), np.zeros((n, 8)), np.zeros((n, 3, 16))
    failed = np.zeros(n, np.uint32)
    lib().qmo_receivers(C.byref(scene.terrain.c), C.byref(scene.obstacles.c), C.byref(scene.weather.c),
                        points.ctypes.data, n, _p(altitude, D), _p(alpha, D), _p(fav, D), _p(failed, U32))
    return dict(altitude_m=altitude, alpha=alpha, favourable=fav, failed=failed)


def evaluate_pairs(scene: abi.Scene, points, sources, pair_point, pair_source, pair_weight=None, threads=None,
                   point_bands=True):
    points = np.ascontiguousarray(points, abi.POINT_DTYPE)
    sources = np.ascontiguousarray(sources, abi.SOURCE_DTYPE)
    pp = np.ascontiguousarray(pair_point, np.uint32)
    ps = np.ascontiguousarray(pair_source, np.uint32)
    pw = None if pair_weight is None else np.ascontiguousarray(pair_weight, np.float32)
    n = pp.size
    sums = np.zeros((n, abi.PAIR_SUMS))
    failed = np.zeros(n, np.uint32)
    bands = np.zeros((points.size, abi.LAYERS, 3, 8)) if point_bands else None
    work = np.zeros(6, np.uint64)
    threads = threads or os.cpu_count() or 1
    rc = lib().qmo_evaluate_pairs(
        C.byref(scene.terrain.c), C.byref(scene.obstacles.c), C.byref(scene.weather.c), points.ctypes.data,
        points.size, sources.ctypes.data, sources.size, _p(pp, U32), _p(ps, U32),
        None if pw is None else _p(pw, F32), n, _p(sums, D), _p(failed, U32),
        None if bands is None else _p(bands, D), threads, _p(work, U64))
    if rc != 0:
        raise ValueError("oracle: pair index out of range")
    keys = ("vertices", "cells", "cells_occupied", "edges", "crossings", "candidates")
    return dict(sums=sums, failed=failed, point_bands=bands, work=dict(zip(keys, (int(w) for w in work))))


def pair_boundaries(scene: abi.Scene, points, sources, pair_point, pair_source, threads=None):
    """Per pair the boundary terms A_ground + A_dif [2 states][8 bands] (dB) and the failure bits."""
    points = np.ascontiguousarray(points, abi.POINT_DTYPE)
    sources = np.ascontiguousarray(sources, abi.SOURCE_DTYPE)
    pp = np.ascontiguousarray(pair_point, np.uint32)
    ps = np.ascontiguousarray(pair_source, np.uint32)
    n = pp.size
    out = np.zeros((n, 2, 8))
    failed = np.zeros(n, np.uint32)
    rc = lib().qmo_pair_boundaries(C.byref(scene.terrain.c), C.byref(scene.obstacles.c), C.byref(scene.weather.c),
                                   points.ctypes.data, points.size, sources.ctypes.data, sources.size, _p(pp, U32),
                                   _p(ps, U32), n, _p(out, D), _p(failed, U32), threads or os.cpu_count() or 1)
    if rc != 0:
        raise ValueError("oracle: pair index out of range")
    return out, failed


def line_nodes(start, end, arcs=(), capacity=4096):
    """The quadrature nodes of a piece (receiver-relative ends), every arc (lo, hi, nearest) handed
    to each wide bucket: dict(along, weight, obstacles, length, subtended, divergence) or None."""
    s = np.ascontiguousarray(start, np.float64)
    e = np.ascontiguousarray(end, np.float64)
    a = np.ascontiguousarray(np.asarray(arcs, np.float64).reshape(-1, 3))
    along, weight = np.zeros(capacity), np.zeros(capacity)
    obst = np.zeros(capacity, np.uint8)
    info = np.zeros(3)
    n = lib().qmo_line_nodes(_p(s, D), _p(e, D), _p(a, D), a.shape[0], capacity, _p(along, D), _p(weight, D),
                             _p(obst, U8), _p(info, D))
    if n < 0:
        return None
    return dict(along=along[:n], weight=weight[:n], obstacles=obst[:n].astype(bool), length=info[0],
                subtended=info[1], divergence=info[2])


def skyline_arcs(obstacles: abi.Obstacles, x, y, lo, hi, radius, floor_m=0.0, capacity=1 << 16):
    out = np.zeros((capacity, 3))
    n = lib().qmo_skyline_arcs(C.byref(obstacles.c), x, y, lo, hi, radius, floor_m, capacity, _p(out, D))
    return out[:min(n, capacity)]


def evaluate_line_pairs(scene: abi.Scene, points, pieces, pair_point, pair_piece, threads=None):
    points = np.ascontiguousarray(points, abi.POINT_DTYPE)
    pieces = np.ascontiguousarray(pieces, abi.LINE_PIECE_DTYPE)
    pp = np.ascontiguousarray(pair_point, np.uint32)
    pl = np.ascontiguousarray(pair_piece, np.uint32)
    n = pp.size
    sums = np.zeros((n, abi.PAIR_SUMS))
    failed = np.zeros(n, np.uint32)
    rays = np.zeros(n, np.uint32)
    rc = lib().qmo_evaluate_line_pairs(C.byref(scene.terrain.c), C.byref(scene.obstacles.c), C.byref(scene.weather.c),
                                       points.ctypes.data, points.size, pieces.ctypes.data, pieces.size, _p(pp, U32),
                                       _p(pl, U32), n, _p(sums, D), _p(failed, U32), _p(rays, U32),
                                       threads or os.cpu_count() or 1)
    if rc != 0:
        raise ValueError("oracle: pair index out of range")
    return dict(sums=sums, failed=failed, rays=rays)


def paint_points(scene: abi.Scene, points, sources, cutoff_m, threads=None):
    points = np.ascontiguousarray(points, abi.POINT_DTYPE)
    sources = np.ascontiguousarray(sources, abi.SOURCE_DTYPE)
    n = points.size
    energy = np.zeros((n, 3))
    paths = np.zeros(n, np.uint64)
    failures = np.zeros(n, np.uint32)
    threads = threads or os.cpu_count() or 1
    lib().qmo_paint_points(C.byref(scene.terrain.c), C.byref(scene.obstacles.c), C.byref(scene.weather.c),
                           points.ctypes.data, n, sources.ctypes.data, sources.size, cutoff_m, _p(energy, D),
                           _p(paths, U64), _p(failures, U32), threads)
    return dict(period_energy=energy, paths=paths, failures=failures)


# ---------------------------------------------------------------------------------------------
# Level distributions and pixel loudness (distribution.cpp).

def kurze_table():
    """The seeded quantile table, (41, 100) float32."""
    t = np.zeros(41 * 100, np.float32)
    lib().qmo_kurze_table(_p(t, F32))
    return t.reshape(41, 100)


def relative_intensity(lam, p):
    return lib().qmo_relative_intensity(float(lam), float(p))


def node_probabilities():
    out = np.zeros((200, 2))
    n = lib().qmo_node_probabilities(_p(out, D), 200)
    return out[:n]


def line_values(kind, x):
    """kind "passes" (mean in progress), "kurze" (lambda) or "event" (duty): (value, share) rows."""
    out = np.zeros((512, 2))
    n = lib().qmo_line_values({"passes": 0, "kurze": 1, "event": 2}[kind], float(x), _p(out, D), 512)
    assert n <= 512
    return out[:n]


def runs(values, shares, scale, intensity, bin_db):
    """Distribution::runs: per bin (shares, weighted), the last row silence; and whether a bin was
    visited twice."""
    v = np.ascontiguousarray(values, np.float64)
    s = np.ascontiguousarray(shares, np.float64)
    bins = int(round(180.0 / bin_db))
    out = np.zeros((bins + 1, 2))
    twice = lib().qmo_runs(_p(v, D), _p(s, D), v.size, float(scale), float(intensity), float(bin_db), _p(out, D))
    return out, bool(twice)


def many_lines_energy(lam, count, bin_db):
    return lib().qmo_many_lines_energy(float(lam), int(count), float(bin_db))


def line_exceeded_db(lam, exceeded, bin_db):
    return lib().qmo_line_exceeded_db(float(lam), float(exceeded), float(bin_db))


def hour_groups():
    """The map's hour groups: slot counts (3, 4) and factors (3 profiles, 3 periods, 4 groups)."""
    slots = np.zeros(12, np.int32)
    factors = np.zeros(36)
    lib().qmo_hour_groups(slots.ctypes.data_as(C.POINTER(C.c_int)), _p(factors, D))
    return slots.reshape(3, 4), factors.reshape(3, 3, 4)


MAP_WEATHER_SPREAD = 10.0 ** 0.15


def fold_pixels(unfolded, offsets, spread=MAP_WEATHER_SPREAD, f64=False):
    """The map's folding (paint/src/sone_fold.rs:115-405 fold_placed_for_map, then
    paint/src/sone.rs:217-234 write_packed) of unfolded pixels (the packed layout in f64) into the
    card's f32 packed pixels at the same offsets, and the floats each takes; with `f64` also the
    folded pixels before the cast (NaN past each pixel's end)."""
    unfolded = np.ascontiguousarray(unfolded, np.float64)
    offsets = np.ascontiguousarray(offsets, np.uint64)
    out = np.full(unfolded.size, np.nan, np.float32)
    out64 = np.full(unfolded.size, np.nan) if f64 else None
    used = np.zeros(offsets.size, np.uint64)
    lib().qmo_fold_pixels(_p(unfolded, D), _p(offsets, U64), offsets.size, float(spread), _p(out, F32),
                          out64.ctypes.data if f64 else None, _p(used, U64))
    return (out, used, out64) if f64 else (out, used)


def merge_close_states(layers, spread=MAP_WEATHER_SPREAD):
    """popup/src/percentiles/weather.rs:139-175 merge_close_states over layers of packed f64 lines
    (a list of (n, 43) arrays); the merged copies."""
    counts = np.array([len(l) for l in layers], np.uint32)
    lines = np.ascontiguousarray(np.concatenate([np.asarray(l, np.float64).reshape(-1, 43) for l in layers]))
    lib().qmo_merge_close_states(_p(lines, D), counts.ctypes.data_as(C.POINTER(C.c_uint32)), counts.size,
                                 float(spread))
    return np.split(lines, np.cumsum(counts)[:-1])


def pixel_loudness(packed, offsets, bin_db=2.0, mixtures=False, threads=None):
    """Per pixel Nden per curve (7: six layers, then the place) and optionally the mixtures
    (pixels, 3 periods, 7 curves, bins, 2)."""
    packed = np.ascontiguousarray(packed, np.float32)
    offsets = np.ascontiguousarray(offsets, np.uint64)
    n = offsets.size
    bins = int(round(180.0 / bin_db))
    nd = np.zeros((n, 7))
    mix = np.zeros((n, 3, 7, bins, 2)) if mixtures else None
    lib().qmo_pixel_loudness(_p(packed, F32), _p(offsets, U64), n, float(bin_db), _p(mix, D) if mixtures else None,
                             _p(nd, D), threads or os.cpu_count() or 1)
    return (nd, mix) if mixtures else nd


def path_difference_loss(d):
    return lib().qmo_path_difference_loss(float(d))


def edge_loss(edge_range, edge_height, source_range, source_height):
    return lib().qmo_edge_loss(float(edge_range), float(edge_height), float(source_range), float(source_height))


def sector_of(east, north, sectors):
    return lib().qmo_sector_of(float(east), float(north), int(sectors))


def horizon_loss(record, sectors, east, north, up):
    r = np.ascontiguousarray(record, np.float32)
    return lib().qmo_horizon_loss(_p(r, F32), int(sectors), float(east), float(north), float(up))


def horizons(scene, x, y, altitude_m, own_footprint=0):
    """Horizons::build at (x, y): QM_HORIZON_FLOATS float32 (terrain's, then buildings') and the
    failure bits."""
    out = np.zeros(abi.HORIZON_FLOATS, np.float32)
    failed = lib().qmo_horizons(C.byref(scene.terrain.c), C.byref(scene.obstacles.c), float(x), float(y),
                                float(altitude_m), int(own_footprint), _p(out, F32))
    return out, failed


BOX_VALUE_FIELDS = (("levels_db", (3, 10)), ("tail_levels_db", (3,)), ("scaled_distance_m", (10,)),
                    ("centroid_m", (3,)), ("axis_rad", ()), ("gradient", ()), ("gradient_spread", ()),
                    ("piece_length_m", (3,)), ("installation_shares", (3, 2, 3)), ("average_length_m", ()))


def npd_distance_levels(classes, track):
    """segment.rs npd_distance_levels of a track's emission: (sel_db[10], scaled_distance_m[10],
    tail_sel_db)."""
    classes = np.ascontiguousarray(classes, abi.AIRCRAFT_CLASS_DTYPE)
    track = np.ascontiguousarray(np.atleast_1d(track), abi.TRACK_DTYPE)
    out = np.zeros(21)
    lib().qmo_npd_distance_levels(classes.ctypes.data, track.ctypes.data, _p(out, D))
    return out[:10], out[10:20], out[20]


def box_values(classes, tracks, period_weights, ends):
    """BoxSums of pieces (track i's emission, its period weights [3], its ends [start xyz, end
    xyz]): a dict of the box values and the average length, or None for a box without energy."""
    classes = np.ascontiguousarray(classes, abi.AIRCRAFT_CLASS_DTYPE)
    tracks = np.ascontiguousarray(tracks, abi.TRACK_DTYPE)
    pw = np.ascontiguousarray(period_weights, np.float64).reshape(-1, 3)
    ends = np.ascontiguousarray(ends, np.float64).reshape(-1, 6)
    assert pw.shape[0] == ends.shape[0] == tracks.size
    out = np.zeros(71)
    if lib().qmo_box_values(classes.ctypes.data, tracks.ctypes.data, _p(pw, D), _p(ends, D), tracks.size,
                            _p(out, D)):
        return None
    values, at = {}, 0
    for name, shape in BOX_VALUE_FIELDS:
        n = int(np.prod(shape)) if shape else 1
        values[name] = out[at:at + n].reshape(shape) if shape else float(out[at])
        at += n
    assert at == 71
    return values


def _pack_values(values):
    return np.concatenate([np.asarray(values[name], np.float64).reshape(-1) for name, _ in BOX_VALUE_FIELDS])


def box_sel(values, centroid, horizon=None):
    """box_sel_at_receiver of box values with the centroid at `centroid` (metres from the
    receiver): SEL energy per period and the installation fractions [3][3]."""
    v = _pack_values(values)
    c = np.ascontiguousarray(centroid, np.float64)
    sel, frac = np.zeros(3), np.zeros(9)
    h = None if horizon is None else np.ascontiguousarray(horizon, np.float32)
    lib().qmo_box_sel(_p(v, D), _p(c, D), None if h is None else _p(h, F32), _p(sel, D), _p(frac, D))
    return sel, frac.reshape(3, 3)


def flight_sums(boxes, points, horizons, segments, n_sums):
    """The card's flight_sums over host records in FP64: [n_sums, FLIGHT_SUMS]."""
    boxes = np.ascontiguousarray(boxes, abi.FLIGHT_BOX_DTYPE)
    points = np.ascontiguousarray(points, abi.FLIGHT_POINT_DTYPE)
    segments = np.ascontiguousarray(segments, abi.FLIGHT_SEGMENT_DTYPE)
    h = np.ascontiguousarray(horizons if horizons is not None else np.zeros(0), np.float32).reshape(-1)
    out = np.zeros((n_sums, abi.FLIGHT_SUMS))
    rc = lib().qmo_flight_sums(boxes.ctypes.data, boxes.size, points.ctypes.data, points.size, _p(h, F32), h.size,
                               segments.ctypes.data, segments.size, _p(out, D), n_sums)
    if rc:
        raise ValueError("a flight record out of range")
    return out
