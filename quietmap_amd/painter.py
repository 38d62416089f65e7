"""The painter's host interface: a Python mirror of the reference's paint-gpu API
(paint-gpu/src/lib.rs:14-22) over libqmpaint's C ABI (include/qm_paint.h).

``Gpu(0)`` opens the card as ``Gpu::new(0)`` does (paint-gpu/src/device.rs:254); ``Gpu.upload``
is ``DeviceSquare::upload`` (paint-gpu/src/square_upload.rs:18); ``Gpu.evaluate_pairs`` is
``Gpu::evaluate_pairs`` (paint-gpu/src/device.rs:440), returning the 9 pair sums and failure bits
per pair and the bands per point; ``Gpu.paint_grid`` paints a receiver grid with every source
within a cutoff. There is no CPU fallback: without the built library or a gfx950 GPU these raise.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import abi

HERE = os.path.dirname(os.path.abspath(__file__))
LIBRARY = os.path.join(HERE, "libqmpaint.so")
# The same library with the path kernels' phase clocks (quietmap_amd/build.py): profiling only.
PHASES_LIBRARY = os.path.join(HERE, "libqmpaint_phases.so")
PHASES = ("outside", "crossings", "setup", "vertex", "profile", "crest", "candidates", "tail", "boundary", "mix")

# Every entry point include/qm_paint.h declares.
EXPORTS = (
    "qm_gpu_new",
    "qm_gpu_name",
    "qm_gpu_free",
    "qm_square_upload",
    "qm_square_bytes",
    "qm_square_free",
    "qm_evaluate_pairs",
    "qm_paint_grid",
    "qm_paint_grid_device",
    "qm_paint_grid_device_noalloc",
    "qm_receivers_device",
    "qm_paint_receivers_device",
    "qm_loudness",
    "qm_loudness_device",
    "qm_aircraft_grid",
    "qm_crossings",
    "qm_evaluate_line_pairs",
    "qm_pixel_loudness",
    "qm_pixel_loudness_unfolded",
    "qm_square_set_lines",
    "qm_square_candidates",
    "qm_evaluate_candidates",
    "qm_kurze_table",
    "qm_horizons",
    "qm_flight_boxes_upload",
    "qm_flight_boxes_count",
    "qm_flight_boxes_free",
    "qm_flight_sums",
    "qm_flight_sums_device",
    "qm_mlp_new",
    "qm_mlp_params",
    "qm_mlp_free",
    "qm_mlp_gradient_device",
    "qm_mlp_train_step_device",
    "qm_mlp_infer_device",
    "qm_surrogate_samples_device",
    "qm_surrogate_paint_receivers_device",
    "qm_level_cells",
    "qm_level_cells_device",
    "qm_hm3_predict",
    "qm_hm3_predict_device",
    "qm_tiles_plan",
    "qm_tiles_run",
    "qm_tiles_paths",
    "qm_tiles_items",
    "qm_tiles_plan_free",
)
RECEIVER_BYTES = 256
NOALLOC_RECEIVERS = 65536  # QM_NOALLOC_RECEIVERS

_lib = None


class PainterError(RuntimeError):
    pass


_phase_build = False


def use_phase_clocks():
    """Loads the phase-clock build instead of libqmpaint.so (before anything else loads it)."""
    global _phase_build
    if _lib is not None and not _phase_build:
        raise PainterError("libqmpaint is already loaded: choose the phase-clock build first")
    _phase_build = True


def phase_clocks(reset: bool = True) -> dict:
    """The phase-clock build's totals of shader-clock cycles per phase of the path (waves' time,
    summed over waves), cleared when `reset`."""
    lib = library()
    if not hasattr(lib, "qm_phase_clocks"):
        raise PainterError("phase clocks need the phase-clock build (painter.use_phase_clocks())")
    out = (C.c_ulonglong * len(PHASES))()
    err = C.create_string_buffer(512)
    _check(lib.qm_phase_clocks(out, int(reset), err, 512), err)
    return dict(zip(PHASES, (int(v) for v in out)))


def library():
    """libqmpaint, loaded once; raises when it has not been built."""
    global _lib
    if _lib is None:
        path = PHASES_LIBRARY if _phase_build else LIBRARY
        if not os.path.exists(path):
            raise PainterError(f"{path} is not built: run __graft_entry__.build() (hipcc, gfx950)")
        lib = C.CDLL(path)
        vp, sz, u32, dbl, cp = C.c_void_p, C.c_size_t, C.c_uint32, C.c_double, C.c_char_p
        lib.qm_gpu_new.argtypes = [C.c_int, C.POINTER(vp), cp, sz]
        lib.qm_gpu_name.argtypes = [vp, cp, sz]
        lib.qm_gpu_free.argtypes = [vp]
        lib.qm_square_upload.argtypes = [vp, vp, vp, vp, vp, sz, dbl, C.POINTER(vp), cp, sz]
        lib.qm_square_bytes.argtypes = [vp]
        lib.qm_square_bytes.restype = sz
        lib.qm_square_free.argtypes = [vp]
        lib.qm_evaluate_pairs.argtypes = [vp, vp, vp, sz, vp, vp, vp, sz, vp, vp, vp, vp, cp, sz]
        lib.qm_paint_grid.argtypes = [vp, vp, vp, u32, u32, dbl, vp, vp, vp, cp, sz]
        lib.qm_paint_grid_device.argtypes = [vp, vp, vp, u32, u32, dbl, vp, vp, vp, vp, cp, sz]
        lib.qm_paint_grid_device_noalloc.argtypes = [vp, vp, vp, u32, u32, dbl, vp, vp, vp, vp, cp, sz]
        lib.qm_receivers_device.argtypes = [vp, vp, vp, sz, vp, vp, cp, sz]
        lib.qm_paint_receivers_device.argtypes = [vp, vp, vp, u32, u32, dbl, vp, vp, vp, vp, cp, sz]
        lib.qm_crossings.argtypes = [vp, vp, vp, sz, u32, vp, vp, vp, vp, vp, vp, vp, cp, sz]
        lib.qm_loudness.argtypes = [vp, vp, sz, vp, cp, sz]
        lib.qm_loudness_device.argtypes = [vp, vp, sz, sz, sz, vp, vp, cp, sz]
        lib.qm_evaluate_line_pairs.argtypes = [vp, vp, vp, sz, vp, sz, vp, vp, sz, vp, vp, vp, vp, vp, cp, sz]
        lib.qm_aircraft_grid.argtypes = [vp, vp, sz, vp, sz, vp, sz, vp, u32, u32, dbl, dbl, vp, vp, vp, vp, vp, vp,
                                         cp, sz]
        lib.qm_pixel_loudness.argtypes = [vp, vp, sz, vp, sz, vp, vp, vp, cp, sz]
        lib.qm_pixel_loudness_unfolded.argtypes = [vp, vp, sz, vp, sz, vp, vp, vp, vp, cp, sz]
        lib.qm_square_set_lines.argtypes = [vp, vp, sz, cp, sz]
        lib.qm_square_candidates.argtypes = [vp]
        lib.qm_square_candidates.restype = sz
        lib.qm_evaluate_candidates.argtypes = [vp, vp, vp, sz, vp, sz, vp, vp, vp, sz, vp, vp, vp, vp, cp, sz]
        lib.qm_kurze_table.argtypes = [vp]
        lib.qm_kurze_table.restype = None
        lib.qm_horizons.argtypes = [vp, vp, vp, vp, sz, vp, vp, cp, sz]
        lib.qm_flight_boxes_upload.argtypes = [vp, vp, sz, C.POINTER(vp), cp, sz]
        lib.qm_flight_boxes_count.argtypes = [vp]
        lib.qm_flight_boxes_count.restype = sz
        lib.qm_flight_boxes_free.argtypes = [vp]
        lib.qm_flight_boxes_free.restype = None
        lib.qm_flight_sums.argtypes = [vp, vp, vp, sz, vp, sz, vp, sz, vp, sz, cp, sz]
        lib.qm_flight_sums_device.argtypes = [vp, vp, vp, vp, vp, sz, vp, vp, cp, sz]
        lib.qm_mlp_new.argtypes = [vp, vp, sz, C.POINTER(vp), cp, sz]
        lib.qm_mlp_params.argtypes = [vp, vp, cp, sz]
        lib.qm_mlp_free.argtypes = [vp]
        lib.qm_mlp_free.restype = None
        lib.qm_mlp_gradient_device.argtypes = [vp, vp, vp, sz, vp, C.POINTER(C.c_float), vp, cp, sz]
        lib.qm_mlp_train_step_device.argtypes = [vp, vp, vp, sz, C.c_float, C.POINTER(C.c_float), vp, cp, sz]
        lib.qm_mlp_infer_device.argtypes = [vp, vp, sz, vp, vp, cp, sz]
        lib.qm_surrogate_samples_device.argtypes = [vp, vp, vp, vp, vp, sz, vp, vp, vp, vp, cp, sz]
        lib.qm_surrogate_paint_receivers_device.argtypes = [vp, vp, vp, vp, u32, u32, dbl, vp, vp, vp, vp, vp, cp, sz]
        lib.qm_level_cells.argtypes = [vp, vp, vp, vp, sz, vp, vp, cp, sz]
        lib.qm_level_cells_device.argtypes = [vp, vp, vp, vp, sz, vp, vp, vp, vp, cp, sz]
        lib.qm_hm3_predict.argtypes = [vp, vp, vp, sz, C.c_int, vp, cp, sz]
        lib.qm_hm3_predict_device.argtypes = [vp, vp, vp, sz, C.c_int, vp, vp, cp, sz]
        lib.qm_tiles_plan.argtypes = [vp, vp, vp, sz, vp, sz, u32, u32, vp, vp, vp, C.POINTER(vp), cp, sz]
        lib.qm_tiles_run.argtypes = [vp, vp, cp, sz]
        lib.qm_tiles_paths.argtypes = [vp]
        lib.qm_tiles_paths.restype = C.c_uint64
        lib.qm_tiles_items.argtypes = [vp]
        lib.qm_tiles_items.restype = C.c_uint64
        lib.qm_tiles_plan_free.argtypes = [vp]
        lib.qm_tiles_plan_free.restype = None
        _lib = lib
    return _lib


def _check(rc, err):
    if rc != 0:
        raise PainterError(err.value.decode(errors="replace"))


def _ptr(a):
    return None if a is None else a.ctypes.data


class Square:
    """A scene on the device (DeviceSquare)."""

    def __init__(self, gpu, handle, scene, sources):
        self.gpu, self.handle, self.scene, self.sources = gpu, handle, scene, sources
        self.lines = np.zeros(0, abi.LINE_PIECE_DTYPE)

    @property
    def bytes(self):
        return library().qm_square_bytes(self.handle)

    def set_lines(self, pieces):
        """The square's line candidates (numbered after its point sources)."""
        pieces = np.ascontiguousarray(pieces, abi.LINE_PIECE_DTYPE)
        err = C.create_string_buffer(512)
        _check(library().qm_square_set_lines(self.handle, _ptr(pieces), pieces.size, err, 512), err)
        self.lines = pieces

    @property
    def candidates(self):
        """Point sources, then line pieces."""
        return int(library().qm_square_candidates(self.handle))

    def layer_of(self, candidate):
        """The layer of each candidate index."""
        c = np.asarray(candidate, np.int64)
        layers = np.concatenate([self.sources["layer"], self.lines["layer"]]).astype(np.int64)
        return layers[c]

    def close(self):
        if self.handle:
            library().qm_square_free(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class FlightBoxes:
    """A square's aircraft boxes on the device (DeviceFlightBoxes)."""

    def __init__(self, handle, n):
        self.handle, self.n = handle, n

    def close(self):
        if self.handle:
            library().qm_flight_boxes_free(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Gpu:
    def __init__(self, ordinal: int = 0):
        lib = library()
        err = C.create_string_buffer(512)
        handle = C.c_void_p()
        _check(lib.qm_gpu_new(ordinal, C.byref(handle), err, 512), err)
        self.handle = handle

    @property
    def name(self) -> str:
        out = C.create_string_buffer(256)
        library().qm_gpu_name(self.handle, out, 256)
        return out.value.decode()

    def close(self):
        if self.handle:
            library().qm_gpu_free(self.handle)
            self.handle = None

    def upload(self, scene: abi.Scene, sources, bucket_m: float = 250.0) -> Square:
        sources = np.ascontiguousarray(sources, abi.SOURCE_DTYPE)
        err = C.create_string_buffer(512)
        handle = C.c_void_p()
        rc = library().qm_square_upload(self.handle, C.byref(scene.terrain.c), C.byref(scene.obstacles.c),
                                        C.byref(scene.weather.c), _ptr(sources), sources.size, bucket_m,
                                        C.byref(handle), err, 512)
        _check(rc, err)
        return Square(self, handle, scene, sources)

    def evaluate_pairs(self, square: Square, points, pair_point, pair_source, pair_weight=None, point_bands=True,
                       pair_bands=False):
        """Pair sums [n, 9], failure bits, point bands [points, 6, 3, 8] and, with `pair_bands`,
        each pair's own mixed bands [n, 3, 8] (Batch::pair_spectra's spectra)."""
        points = np.ascontiguousarray(points, abi.POINT_DTYPE)
        pp = np.ascontiguousarray(pair_point, np.uint32)
        ps = np.ascontiguousarray(pair_source, np.uint32)
        pw = None if pair_weight is None else np.ascontiguousarray(pair_weight, np.float32)
        n = pp.size
        sums = np.zeros((n, abi.PAIR_SUMS), np.float32)
        failed = np.zeros(n, np.uint32)
        bands = np.zeros((points.size, abi.LAYERS, abi.PERIODS, abi.BANDS), np.float32) if point_bands else None
        spectra = np.zeros((n, abi.PERIODS, abi.BANDS), np.float32) if pair_bands else None
        err = C.create_string_buffer(512)
        rc = library().qm_evaluate_pairs(self.handle, square.handle, _ptr(points), points.size, _ptr(pp), _ptr(ps),
                                         _ptr(pw), n, _ptr(sums), _ptr(failed), _ptr(bands), _ptr(spectra), err, 512)
        _check(rc, err)
        return dict(sums=sums, failed=failed, point_bands=bands, pair_bands=spectra)

    def paint_grid(self, square: Square, points, nx: int, ny: int, cutoff_m: float):
        points = np.ascontiguousarray(points, abi.POINT_DTYPE)
        assert points.size == nx * ny
        energy = np.zeros((points.size, abi.PERIODS), np.float64)
        paths = np.zeros(points.size, np.uint32)
        failures = np.zeros(points.size, np.uint32)
        err = C.create_string_buffer(512)
        rc = library().qm_paint_grid(self.handle, square.handle, _ptr(points), nx, ny, cutoff_m, _ptr(energy),
                                     _ptr(paths), _ptr(failures), err, 512)
        _check(rc, err)
        return dict(period_energy=energy, paths=paths, failures=failures)

    def paint_grid_device(self, square: Square, d_points: int, nx: int, ny: int, cutoff_m: float, d_energy: int,
                          d_paths: int, d_failures: int, stream: int = 0):
        """Device pointers in and out (inputs resident in HBM); enqueued on `stream`."""
        err = C.create_string_buffer(512)
        rc = library().qm_paint_grid_device(self.handle, square.handle, d_points, nx, ny, cutoff_m, d_energy,
                                            d_paths, d_failures, stream or None, err, 512)
        _check(rc, err)

    def paint_grid_device_noalloc(self, square: Square, d_points: int, nx: int, ny: int, cutoff_m: float,
                                  d_energy: int, d_paths: int, d_failures: int, stream: int = 0):
        """paint_grid_device without device allocations (the library's static scratch; at most
        NOALLOC_RECEIVERS receivers; calls overlapping on one device race)."""
        err = C.create_string_buffer(512)
        rc = library().qm_paint_grid_device_noalloc(self.handle, square.handle, d_points, nx, ny, cutoff_m, d_energy,
                                                    d_paths, d_failures, stream or None, err, 512)
        _check(rc, err)

    def receivers_device(self, square: Square, d_points: int, n: int, d_receivers: int, stream: int = 0):
        """Receiver properties of `n` device points into `d_receivers` (256 bytes each)."""
        err = C.create_string_buffer(512)
        rc = library().qm_receivers_device(self.handle, square.handle, d_points, n, d_receivers, stream or None,
                                           err, 512)
        _check(rc, err)

    def receivers(self, square: Square, points):
        """The receiver properties (abi.RECEIVER_DTYPE) of host points, computed on the card."""
        from . import hip

        points = np.ascontiguousarray(points, abi.POINT_DTYPE)
        if points.size == 0:
            return np.zeros(0, abi.RECEIVER_DTYPE)
        d_points = hip.DeviceBuffer.from_array(points)
        d_receivers = hip.DeviceBuffer(points.size * RECEIVER_BYTES)
        self.receivers_device(square, d_points.ptr, points.size, d_receivers.ptr)
        hip.synchronize()
        return d_receivers.to_array(abi.RECEIVER_DTYPE, points.size)

    def paint_receivers_device(self, square: Square, d_receivers: int, nx: int, ny: int, cutoff_m: float,
                               d_energy: int, d_paths: int, d_failures: int, stream: int = 0):
        err = C.create_string_buffer(512)
        rc = library().qm_paint_receivers_device(self.handle, square.handle, d_receivers, nx, ny, cutoff_m,
                                                 d_energy, d_paths, d_failures, stream or None, err, 512)
        _check(rc, err)

    def loudness(self, levels):
        """Zwicker loudness (sone) of third-octave levels (n, 28) dB."""
        lv = np.ascontiguousarray(np.asarray(levels, np.float32).reshape(-1, 28))
        out = np.zeros(lv.shape[0], np.float32)
        err = C.create_string_buffer(512)
        _check(library().qm_loudness(self.handle, _ptr(lv), lv.shape[0], _ptr(out), err, 512), err)
        return out

    def loudness_device(self, d_levels: int, n: int, cell_stride: int, band_stride: int, d_sone: int,
                        stream: int = 0):
        err = C.create_string_buffer(512)
        rc = library().qm_loudness_device(self.handle, d_levels, n, cell_stride, band_stride, d_sone,
                                          stream or None, err, 512)
        _check(rc, err)

    def aircraft_grid(self, scene, passes=False, screening=None):
        """Doc 29 aircraft outputs of an AircraftScene's receiver grid: SEL energy per period,
        N65, N70 and the LAmax histogram per receiver, with `passes` the pass classes
        [n, 2, 3, 48] (energies, then passing shares, per period and class), and with `screening`
        (a dict of `horizons` [records, QM_HORIZON_FLOATS] float32, `horizon_of` per receiver and
        optional `vertex_ground_m` per vertex) the SEL screened by the receivers' horizons."""
        nx, ny = scene.grid_shape
        xy = np.ascontiguousarray(scene.points_xy, np.float64).reshape(-1, 2)
        assert xy.shape[0] == nx * ny
        classes = np.ascontiguousarray(scene.classes, abi.AIRCRAFT_CLASS_DTYPE)
        tracks = np.ascontiguousarray(scene.tracks, abi.TRACK_DTYPE)
        verts = np.ascontiguousarray(scene.vertices, np.float32).reshape(-1, 3)
        n = nx * ny
        energy = np.zeros((n, 3))
        n65, n70 = np.zeros(n, np.uint32), np.zeros(n, np.uint32)
        hist = np.zeros((n, abi.HIST_BINS), np.uint32)
        pc = np.zeros((n, 2, 3, abi.PASS_CLASSES)) if passes else None
        scr = None
        if screening is not None:
            h = np.ascontiguousarray(screening["horizons"], np.float32).reshape(-1, abi.HORIZON_FLOATS)
            of = np.ascontiguousarray(screening["horizon_of"], np.uint32)
            assert of.size == n
            gr = screening.get("vertex_ground_m")
            gr = None if gr is None else np.ascontiguousarray(gr, np.float32)
            assert gr is None or gr.size == verts.shape[0]
            scr = abi.AircraftScreening(h.ctypes.data, h.shape[0], of.ctypes.data, None if gr is None else gr.ctypes.data)
        err = C.create_string_buffer(512)
        rc = library().qm_aircraft_grid(self.handle, _ptr(classes), classes.size, _ptr(tracks), tracks.size,
                                        _ptr(verts), verts.shape[0], _ptr(xy), nx, ny, scene.receiver_altitude_m,
                                        scene.reach_m, _ptr(energy), _ptr(n65), _ptr(n70), _ptr(hist),
                                        _ptr(pc) if passes else None, C.byref(scr) if scr is not None else None,
                                        err, 512)
        _check(rc, err)
        out = dict(period_energy=energy, n65=n65, n70=n70, hist=hist)
        if passes:
            out["passes"] = pc
        return out

    def pixel_loudness(self, packed, offsets, kurze=None, mixtures=False):
        """Nden per curve (pixels, 7: each layer alone, then the place) of packed pixels
        (quietmap_amd.pixels), and with `mixtures` their per-period mixtures
        (pixels, 3, 7, 90, 2)."""
        packed = np.ascontiguousarray(packed, np.float32)
        offsets = np.ascontiguousarray(offsets, np.uint64)
        n = offsets.size
        nd = np.zeros((n, 7), np.float32)
        mix = np.zeros((n, 3, 7, 90, 2), np.float32) if mixtures else None
        table = None if kurze is None else np.ascontiguousarray(kurze, np.float32)
        err = C.create_string_buffer(512)
        rc = library().qm_pixel_loudness(self.handle, _ptr(packed), packed.size, _ptr(offsets), n,
                                         _ptr(table) if table is not None else None,
                                         _ptr(mix) if mixtures else None, _ptr(nd), err, 512)
        _check(rc, err)
        return (nd, mix) if mixtures else nd

    def evaluate_candidates(self, square: Square, receivers, candidate, receiver_of_pair, point_of_receiver=None,
                            n_points=None, band_weight=None, bands=None, pair_spectra=False):
        """Gpu::evaluate_pairs as GpuBatch calls it: pairs of point and line candidates of one
        square, each pair's receiver among `receivers`, each receiver's point in `bands` (an
        array [n_points, 6, 3, 8] float32 added to; None: no bands). Returns a dict of `sums`
        [n, 9], `failed` [n] and, with `pair_spectra`, `pair_spectra` [n, 3, 8]."""
        rcv = np.ascontiguousarray(receivers, abi.POINT_DTYPE)
        cand = np.ascontiguousarray(candidate, np.uint32)
        rop = np.ascontiguousarray(receiver_of_pair, np.uint32)
        n = cand.size
        assert rop.size == n
        por = None if point_of_receiver is None else np.ascontiguousarray(point_of_receiver, np.uint32)
        assert por is None or por.size == rcv.size
        w = None if band_weight is None else np.ascontiguousarray(band_weight, np.float32)
        assert w is None or w.size == n
        if bands is not None:
            assert bands.dtype == np.float32 and bands.flags["C_CONTIGUOUS"]
            n_points = bands.size // (abi.LAYERS * abi.PERIODS * abi.BANDS)
        sums = np.zeros((n, abi.PAIR_SUMS), np.float32)
        failed = np.zeros(n, np.uint32)
        spectra = np.zeros((n, abi.PERIODS, abi.BANDS), np.float32) if pair_spectra else None
        err = C.create_string_buffer(512)
        rc = library().qm_evaluate_candidates(self.handle, square.handle, _ptr(rcv), rcv.size, _ptr(por),
                                              int(n_points or 0), _ptr(cand), _ptr(rop), _ptr(w), n, _ptr(sums),
                                              _ptr(failed), _ptr(bands), _ptr(spectra), err, 512)
        _check(rc, err)
        out = dict(sums=sums, failed=failed)
        if pair_spectra:
            out["pair_spectra"] = spectra
        return out

    def pixel_loudness_unfolded(self, unfolded, offsets, kurze=None, mixtures=False, folded=False):
        """Nden per curve (pixels, 7) of unfolded pixels (the packed layout in f64, lines as placed),
        folded for the map on the card first; with `mixtures` the per-period mixtures and with
        `folded` the folded f32 pixels (NaN past each pixel's end) too, in that order."""
        unfolded = np.ascontiguousarray(unfolded, np.float64)
        offsets = np.ascontiguousarray(offsets, np.uint64)
        n = offsets.size
        nd = np.zeros((n, 7), np.float32)
        mix = np.zeros((n, 3, 7, 90, 2), np.float32) if mixtures else None
        out = np.zeros(unfolded.size, np.float32) if folded else None
        table = None if kurze is None else np.ascontiguousarray(kurze, np.float32)
        err = C.create_string_buffer(512)
        rc = library().qm_pixel_loudness_unfolded(self.handle, _ptr(unfolded), unfolded.size, _ptr(offsets), n,
                                                  _ptr(table) if table is not None else None,
                                                  _ptr(mix) if mixtures else None, _ptr(nd),
                                                  _ptr(out) if folded else None, err, 512)
        _check(rc, err)
        result = (nd,) + ((mix,) if mixtures else ()) + ((out,) if folded else ())
        return result[0] if len(result) == 1 else result

    def horizons(self, square: Square, points, altitude_m=None):
        """Terrain and building horizons of `points` (Horizons::build, write_f32): floats
        [n, QM_HORIZON_FLOATS] and failure bits per point; at `altitude_m` per point, or the
        receiver altitude (ground + 4 m) when None."""
        pts = np.ascontiguousarray(points, abi.POINT_DTYPE)
        n = pts.size
        alt = None if altitude_m is None else np.ascontiguousarray(altitude_m, np.float64)
        assert alt is None or alt.size == n
        out = np.zeros((n, abi.HORIZON_FLOATS), np.float32)
        failed = np.zeros(n, np.uint32)
        err = C.create_string_buffer(512)
        rc = library().qm_horizons(self.handle, square.handle, _ptr(pts), _ptr(alt), n, _ptr(out), _ptr(failed),
                                   err, 512)
        _check(rc, err)
        return out, failed

    def flight_boxes(self, boxes):
        """Uploads aircraft box records (abi.FLIGHT_BOX_DTYPE): a FlightBoxes handle."""
        boxes = np.ascontiguousarray(boxes, abi.FLIGHT_BOX_DTYPE)
        err = C.create_string_buffer(512)
        handle = C.c_void_p()
        _check(library().qm_flight_boxes_upload(self.handle, _ptr(boxes), boxes.size, C.byref(handle), err, 512), err)
        return FlightBoxes(handle, boxes.size)

    def flight_sums(self, boxes, points, horizons, segments, n_sums):
        """Per sum slot the boxes' energy per period and pass classes [n_sums, FLIGHT_SUMS]."""
        pts = np.ascontiguousarray(points, abi.FLIGHT_POINT_DTYPE)
        segs = np.ascontiguousarray(segments, abi.FLIGHT_SEGMENT_DTYPE)
        h = None if horizons is None else np.ascontiguousarray(horizons, np.float32).reshape(-1)
        out = np.zeros((n_sums, abi.FLIGHT_SUMS))
        err = C.create_string_buffer(512)
        rc = library().qm_flight_sums(self.handle, boxes.handle, _ptr(pts), pts.size, _ptr(h), 0 if h is None else h.size,
                                      _ptr(segs), segs.size, _ptr(out), n_sums, err, 512)
        _check(rc, err)
        return out

    def flight_sums_device(self, boxes, d_points: int, d_horizons: int, d_segments: int, n_segments: int,
                           d_sums: int, stream: int = 0):
        err = C.create_string_buffer(512)
        rc = library().qm_flight_sums_device(self.handle, boxes.handle, d_points, d_horizons or None, d_segments,
                                             n_segments, d_sums, stream or None, err, 512)
        _check(rc, err)

    def evaluate_line_pairs(self, square: Square, points, pieces, pair_point, pair_piece, point_bands=False,
                            pair_bands=False):
        """Line pairs (point, piece): sums [n, 9], failure bits, rays per pair, point bands and pair
        bands."""
        pts = np.ascontiguousarray(points, abi.POINT_DTYPE)
        pcs = np.ascontiguousarray(pieces, abi.LINE_PIECE_DTYPE)
        pp = np.ascontiguousarray(pair_point, np.uint32)
        pl = np.ascontiguousarray(pair_piece, np.uint32)
        n = pp.size
        sums = np.zeros((n, abi.PAIR_SUMS), np.float32)
        failed = np.zeros(n, np.uint32)
        rays = np.zeros(n, np.uint32)
        bands = np.zeros((pts.size, abi.LAYERS, 3, 8), np.float32) if point_bands else None
        spectra = np.zeros((n, 3, 8), np.float32) if pair_bands else None
        err = C.create_string_buffer(512)
        rc = library().qm_evaluate_line_pairs(self.handle, square.handle, _ptr(pts), pts.size, _ptr(pcs), pcs.size,
                                              _ptr(pp), _ptr(pl), n, _ptr(sums), _ptr(failed), _ptr(rays),
                                              _ptr(bands), _ptr(spectra), err, 512)
        _check(rc, err)
        return dict(sums=sums, failed=failed, rays=rays, point_bands=bands, pair_bands=spectra)

    def level_cells(self, assessed, layer_energy, ndens):
        """PaintedCells::from_painted: per pixel its Nden and Lden cells (pixels, 7) uint8 each."""
        assessed = np.ascontiguousarray(assessed, np.uint8).reshape(-1)
        n = assessed.size
        energy = np.ascontiguousarray(layer_energy, np.float64).reshape(n, 6)
        nd = np.ascontiguousarray(ndens, np.float64).reshape(n, 7)
        nden = np.zeros((n, 7), np.uint8)
        lden = np.zeros((n, 7), np.uint8)
        err = C.create_string_buffer(512)
        rc = library().qm_level_cells(self.handle, _ptr(assessed), _ptr(energy), _ptr(nd), n, _ptr(nden), _ptr(lden),
                                      err, 512)
        _check(rc, err)
        return nden, lden

    def hm3_predict(self, cells, tile_layer, metric="nden"):
        """encode_metric_tile before Brotli for tiles of 512 x 512 cells: (tiles, 6 + 512 * 512) bytes."""
        cells = np.ascontiguousarray(cells, np.uint8).reshape(-1, 512 * 512)
        layers = np.ascontiguousarray(tile_layer, np.uint8).reshape(-1)
        assert layers.size == cells.shape[0]
        out = np.zeros((cells.shape[0], 6 + 512 * 512), np.uint8)
        err = C.create_string_buffer(512)
        rc = library().qm_hm3_predict(self.handle, _ptr(cells), _ptr(layers), layers.size,
                                      0 if metric == "nden" else 1, _ptr(out), err, 512)
        _check(rc, err)
        return out

    def crossings(self, square: Square, segments, capacity: int = 48):
        seg = np.ascontiguousarray(segments, np.float64).reshape(-1, 4)
        n = seg.shape[0]
        t = np.zeros((n, capacity), np.float64)
        fp = np.zeros((n, capacity), np.uint64)
        ol = np.zeros((n, capacity), np.uint32)
        ed = np.zeros((n, capacity), np.uint32)
        fl = np.zeros((n, capacity), np.uint32)
        count = np.zeros(n, np.uint32)
        failed = np.zeros(n, np.uint32)
        err = C.create_string_buffer(512)
        rc = library().qm_crossings(self.handle, square.handle, _ptr(seg), n, capacity, _ptr(t), _ptr(fp), _ptr(ol),
                                    _ptr(ed), _ptr(fl), _ptr(count), _ptr(failed), err, 512)
        _check(rc, err)
        return dict(t=t, footprint=fp, outline=ol, edge=ed, flags=fl, count=count, failed=failed)


def lden_db(period_energy):
    """Lden of per-period energies (physics/src/bands.rs:59-66), dB; -inf for silence."""
    e = np.asarray(period_energy, np.float64)
    hours = np.array([12.0, 4.0, 8.0])
    penalty = 10.0 ** (np.array([0.0, 5.0, 10.0]) / 10.0)
    total = (e * hours * penalty).sum(axis=-1) / 24.0
    with np.errstate(divide="ignore"):
        return 10.0 * np.log10(total)


def kurze_table():
    """The library's quantile table of Kurze's line statistics, (41, 100) float32 (host code)."""
    t = np.zeros(41 * 100, np.float32)
    library().qm_kurze_table(_ptr(t))
    return t.reshape(41, 100)
