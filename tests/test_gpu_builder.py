"""GPU octree builder and volume generators against the CPU oracle."""
import numpy as np
import pytest

from conftest import blobs_volume

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("shape,iso", [((16, 16, 16), 0.5), ((13, 16, 9), 0.3), ((64, 64, 64), 0.4),
                                       ((100, 128, 77), 0.25), ((32, 32, 32), 0.0)])
@pytest.mark.parametrize("literal", [False, True])
def test_builder_runs_match_oracle(vt, vto, ctx, tmp_path, shape, iso, literal):
    vol = blobs_volume(max(shape), 10, seed=sum(shape), dims=shape)
    if iso == 0.0:
        vol = vol - 0.3   # a field with a real zero crossing, and zero padding (Q11)
    print(f"builder {shape} iso={iso} literal={literal}", flush=True)
    n_c, runs_c = vto.build_octree(vol, iso, literal=literal)
    d = ctx.to_device(vol)
    n_g, sizes = ctx.octree_build(d, shape, iso, literal, min(3, n_c))
    assert n_g == n_c
    for l in range(n_c + 1):
        got = ctx.octree_runs(l, sizes[l])
        assert np.array_equal(got, runs_c[l]), f"level {l}"
    # identical reference-format files
    fg, fc = tmp_path / "g.oct", tmp_path / "c.oct"
    ctx.octree_write_file(fg)
    vto.octree_write(fc, iso, 1 << n_c, shape, n_c, runs_c)
    assert fg.read_bytes() == fc.read_bytes()
    d.free()


def test_generators_deterministic_and_sliced(vt, ctx):
    dims = (48, 40, 36)
    nvox = int(np.prod(dims))
    full, part = ctx.alloc(nvox * 4), ctx.alloc(nvox * 4)
    for kind, kw in [(vt.GEN_FBM, dict(octaves=6, base_freq=1 / 16, gain=0.5, lacunarity=2.0, scale=1.0)),
                     (vt.GEN_BLOBS, dict(shapes=np.array([[0.5, 0.5, 0.5, 0.1, 1.0], [0.2, 0.7, 0.3, 0.05, 0.8]],
                                                        np.float32), n=48)),
                     (vt.GEN_SDF, dict(shapes=np.array([[0, 20, 20, 18, 8, 0, 0, 0], [1, 30, 10, 20, 9, 3, 1, 0]],
                                                       np.float32), n=48, octaves=2, base_freq=1 / 8, gain=0.5,
                                       lacunarity=2.0, scale=0.02)),
                     (vt.GEN_FOURIER, dict(shapes=np.array([[0.1, 0.2, 0.05, 1.0, 1.0], [0.3, -0.1, 0.2, 0.5, 0.3]],
                                                           np.float32)))]:
        ctx.volume_generate(kind, dims, 1234, full, **kw)
        a = full.download(np.float32, dims)
        # two x-slabs generate the same values as one call
        ctx.volume_generate(kind, dims, 1234, part, x0=0, nx=20, **kw)
        b1 = part.download(np.float32, (20,) + dims[1:])
        ctx.volume_generate(kind, dims, 1234, part, x0=20, nx=dims[0] - 20, **kw)
        b2 = part.download(np.float32, (dims[0] - 20,) + dims[1:])
        assert np.array_equal(a, np.concatenate([b1, b2])), kind
        assert np.isfinite(a).all()
        if kind == vt.GEN_FBM:
            assert -1.0 <= a.min() < -0.2 and 0.2 < a.max() <= 1.0
            ctx.volume_generate(kind, dims, 1235, part, **kw)
            assert not np.array_equal(a, part.download(np.float32, dims))
    full.free()
    part.free()


def test_volume_write_hdf5_roundtrip(vt, ctx, tmp_path):
    dims = (70, 33, 20)
    vol = np.random.default_rng(5).standard_normal(dims).astype(np.float32)
    d = ctx.to_device(vol)
    for chunk in (0, 16):
        p = tmp_path / f"w{chunk}.h5"
        ctx.volume_write_hdf5(d, dims, p, "value", chunk)
        with vt.File(p, "value") as f:
            assert f.dims == dims
            assert np.array_equal(f.read_box((0, 0, 0), dims), vol)
    d.free()
