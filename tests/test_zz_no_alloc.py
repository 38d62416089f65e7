"""The allocation-free grid painter (qm_paint_grid_device_noalloc, no counterpart in the
reference) against the allocating one (qm_paint_grid_device) on the same device inputs, called
through the C ABI: two calls in a row, the second on a smaller grid, so that the second reads the
static scratch the first left behind. Per receiver the painter sums its sources in one fixed order,
so energies, path counts and failure counts agree bit for bit."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gpu():
    from quietmap_amd.painter import Gpu

    g = Gpu(0)
    yield g
    g.close()


def paint_both(gpu, square, points, nx, ny, cutoff_m, stream):
    from quietmap_amd.hip import DeviceBuffer, synchronize

    n = nx * ny
    d_points = DeviceBuffer.from_array(points)
    out = {}
    for which in ("alloc", "noalloc"):
        d_energy, d_paths, d_fail = DeviceBuffer(8 * 3 * n), DeviceBuffer(4 * n), DeviceBuffer(4 * n)
        call = gpu.paint_grid_device if which == "alloc" else gpu.paint_grid_device_noalloc
        call(square, d_points.ptr, nx, ny, cutoff_m, d_energy.ptr, d_paths.ptr, d_fail.ptr, stream)
        synchronize()
        out[which] = (d_energy.to_array(np.float64, (n, 3)), d_paths.to_array(np.uint32, n),
                      d_fail.to_array(np.uint32, n))
        for b in (d_energy, d_paths, d_fail):
            b.close()
    d_points.close()
    return out


def test_noalloc_painter_equals_the_allocating_one_twice(gpu):
    from quietmap_amd import scenes
    from quietmap_amd.hip import Stream

    sc = scenes.tile_scene(scale=0.05)
    nx, ny = sc.grid_shape
    print(f"case: no-alloc painter, tile scene {nx}x{ny}", flush=True)
    square = gpu.upload(sc.scene, sc.sources)
    grid = sc.points.reshape(ny, nx)
    stream = Stream()
    for (gx, gy) in ((min(nx, 200), min(ny, 160)), (37, 23)):
        assert gx * gy <= 65536
        pts = np.ascontiguousarray(grid[:gy, :gx]).reshape(-1)
        out = paint_both(gpu, square, pts, gx, gy, sc.cutoff_m, stream.handle)
        (ea, pa, fa), (eb, pb, fb) = out["alloc"], out["noalloc"]
        print(f"grid {gx}x{gy}: paths {int(pa.sum())}, failures {int(fa.sum())}", flush=True)
        assert pa.sum() > 0
        assert np.array_equal(pa, pb) and np.array_equal(fa, fb)
        assert ea.tobytes() == eb.tobytes()
    stream.close()
    square.close()


def test_noalloc_painter_refuses_more_than_its_receivers(gpu):
    from quietmap_amd import scenes
    from quietmap_amd.painter import PainterError

    sc = scenes.tile_scene(scale=0.02)
    square = gpu.upload(sc.scene, sc.sources)
    with pytest.raises(PainterError, match="exceed the allocation-free painter"):
        # Rejected on the host before any device pointer is touched.
        gpu.paint_grid_device_noalloc(square, 1, 300, 300, sc.cutoff_m, 1, 1, 1)
    square.close()
