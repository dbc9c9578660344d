"""MRC2014 reader/writer and the file-driver / command-line plumbing, on the
CPU. The format is pinned by bytes placed at the MRC2014 header offsets
(Cheng et al. 2015) and the FEI metadata-block offsets with `struct`,
independently of the module's own structured types."""
import struct

import numpy
import pytest

from guanaco_amd import mrc


def spec_header(nx, ny, nz, mode, mx, my, mz, cell, ispg, nsymbt,
                exttyp=b"\0\0\0\0", order="<", machst=None):
    """A 1024-byte MRC2014 header built field by field at its offsets."""
    h = bytearray(1024)
    struct.pack_into(order + "4i", h, 0, nx, ny, nz, mode)
    struct.pack_into(order + "3i", h, 28, mx, my, mz)
    struct.pack_into(order + "3f", h, 40, *cell)
    struct.pack_into(order + "3f", h, 52, 90.0, 90.0, 90.0)
    struct.pack_into(order + "3i", h, 64, 1, 2, 3)
    struct.pack_into(order + "2i", h, 88, ispg, nsymbt)
    h[104:108] = exttyp
    struct.pack_into(order + "i", h, 108, 20140)
    h[208:212] = b"MAP "
    if machst is None:
        machst = b"\x44\x44\0\0" if order == "<" else b"\x11\x11\0\0"
    h[212:216] = machst
    return bytes(h)


def fei_block(size, tilt, defocus, order="<"):
    b = bytearray(size)
    struct.pack_into(order + "i", b, 0, size)
    struct.pack_into(order + "d", b, 100, tilt)
    struct.pack_into(order + "d", b, 220, defocus)
    return bytes(b)


@pytest.mark.parametrize("order", ["<", ">"])
def test_reads_a_spec_built_tilt_series(tmp_path, order):
    T, Y, X, size = 3, 4, 5, 800   # an FEI block larger than FEI1's 768
    data = numpy.arange(T * Y * X, dtype=numpy.float32).reshape(T, Y, X)
    ext = b"".join(fei_block(size, t, 1e4 + t, order) for t in (-30.0, 0.5,
                                                                 30.0))
    raw = spec_header(X, Y, T, 2, X, Y, 1, (1.5 * X, 1.5 * Y, 1.5), 0,
                      len(ext), b"FEI1", order)
    p = tmp_path / "tilts.mrc"
    p.write_bytes(raw + ext + data.astype(order + "f4").tobytes())
    with mrc.open(p) as f:
        assert f.data.shape == (T, Y, X) and f.is_image_stack()
        numpy.testing.assert_array_equal(f.data, data)
        v = f.voxel_size
        assert (v["x"], v["y"], v["z"]) == (1.5, 1.5, 1.5)
        h = f.indexed_extended_header
        assert h.shape == (T,) and h.dtype.itemsize == size
        numpy.testing.assert_array_equal(h["Alpha tilt"], [-30.0, 0.5, 30.0])
        numpy.testing.assert_array_equal(h["Defocus"],
                                         [1e4 - 30, 1e4 + 0.5, 1e4 + 30])
        assert (h["Metadata size"] == size).all()


def test_machine_stamp_missing_falls_back_on_the_mode(tmp_path):
    data = numpy.ones((2, 3), numpy.int16)
    raw = spec_header(3, 2, 1, 1, 3, 2, 1, (3, 2, 1), 0, 0, order=">",
                      machst=b"\0\0\0\0")
    p = tmp_path / "img.mrc"
    p.write_bytes(raw + data.astype(">i2").tobytes())
    with mrc.open(p) as f:
        assert f.data.shape == (2, 3) and f.data.dtype == numpy.dtype(">i2")
        numpy.testing.assert_array_equal(f.data, data)


def test_written_file_has_spec_layout(tmp_path):
    T, ndf, Y, X = 2, 3, 4, 5
    data = numpy.random.default_rng(0).random((T, ndf, Y, X),
                                              dtype=numpy.float32)
    ext = numpy.zeros((T, ndf), mrc.FEI1_EXTENDED_HEADER_DTYPE)
    ext["Metadata size"] = 768
    ext["Alpha tilt"] = [[-10.0] * ndf, [20.0] * ndf]
    ext["Defocus"] = numpy.arange(ndf) * 100.0 - 100.0
    p = tmp_path / "stack.mrc"
    mrc.write(p, data, voxel_size=(1.5, 1.5, 2.0), extended_header=ext,
              exttyp=b"FEI1")
    raw = p.read_bytes()
    nx, ny, nz, mode = struct.unpack_from("<4i", raw, 0)
    mx, my, mz = struct.unpack_from("<3i", raw, 28)
    cell = struct.unpack_from("<3f", raw, 40)
    ispg, nsymbt = struct.unpack_from("<2i", raw, 88)
    assert (nx, ny, nz, mode) == (X, Y, T * ndf, 2)
    assert (mx, my, mz) == (X, Y, ndf) and ispg == 401
    assert cell == (1.5 * X, 1.5 * Y, 2.0 * ndf)
    assert struct.unpack_from("<3i", raw, 64) == (1, 2, 3)
    assert nsymbt == T * ndf * 768 and raw[104:108] == b"FEI1"
    assert struct.unpack_from("<i", raw, 108)[0] == 20140
    assert raw[208:212] == b"MAP " and raw[212:214] == b"\x44\x44"
    dmin, dmax, dmean = struct.unpack_from("<3f", raw, 76)
    rms = struct.unpack_from("<f", raw, 216)[0]
    assert numpy.isclose(dmin, data.min()) and numpy.isclose(dmax, data.max())
    assert numpy.isclose(dmean, data.mean(dtype=numpy.float64))
    assert numpy.isclose(rms, data.std(dtype=numpy.float64), rtol=1e-6)
    assert len(raw) == 1024 + nsymbt + data.nbytes
    blk = raw[1024 + 768 * 4:1024 + 768 * 5]   # tilt 1, defocus 1
    assert struct.unpack_from("<i", blk, 0)[0] == 768
    assert struct.unpack_from("<d", blk, 100)[0] == 20.0
    assert struct.unpack_from("<d", blk, 220)[0] == 0.0
    body = numpy.frombuffer(raw[1024 + nsymbt:], "<f4").reshape(data.shape)
    numpy.testing.assert_array_equal(body, data)


def test_fei_offsets_follow_the_field_sizes():
    # each field starts where the previous one ends, except the 48 unused
    # bytes after "Pixel size Y"; a mistyped offset breaks the chain
    dt = mrc.FEI1_EXTENDED_HEADER_DTYPE
    names = list(dt.names)
    for a, b in zip(names, names[1:]):
        end = dt.fields[a][1] + dt.fields[a][0].itemsize
        gap = 48 if a == "Pixel size Y" else 0
        assert dt.fields[b][1] == end + gap, (a, b)
    assert dt.fields["Alpha tilt"][1] == 100
    assert dt.fields["Defocus"][1] == 220 and dt.itemsize == 768


@pytest.mark.parametrize("shape,ispg,nz,mz,dtype", [
    ((6, 7), 0, 1, 1, numpy.int8),
    ((4, 6, 7), 1, 4, 4, numpy.float32),
    ((2, 3, 6, 7), 401, 6, 3, numpy.float16),
    ((3, 6, 7), 1, 3, 3, numpy.complex64),
    ((3, 6, 7), 1, 3, 3, numpy.uint16),
])
def test_new_files_shapes_and_modes(tmp_path, shape, ispg, nz, mz, dtype):
    p = tmp_path / "x.mrc"
    data = (numpy.arange(numpy.prod(shape)) % 100).reshape(shape).astype(
        dtype)
    with mrc.new_mmap(p, shape, mrc.mode_from_dtype(dtype)) as f:
        assert (f.header["dmax"] < f.header["dmin"]) and f.header["rms"] < 0
        f.data[...] = data
        f.voxel_size = 2.0
    with mrc.open(p) as f:
        assert f.data.shape == shape and int(f.header["ispg"]) == ispg
        assert (int(f.header["nz"]), int(f.header["mz"])) == (nz, mz)
        numpy.testing.assert_array_equal(f.data, data)
        assert float(f.voxel_size["z"]) == 2.0


def test_image_stack_and_read_only(tmp_path):
    p = tmp_path / "s.mrc"
    with mrc.new_mmap(p, (5, 4, 3)) as f:
        f.set_image_stack()
        f.data[2] = 7
    with mrc.open(p) as f:
        assert f.is_image_stack() and int(f.header["mz"]) == 1
        assert f.data.shape == (5, 4, 3) and f.data[2].sum() == 7 * 12
        with pytest.raises(ValueError, match="read-only"):
            f.voxel_size = 1.0
    with pytest.raises(ValueError, match="exists"):
        mrc.new_mmap(p, (2, 2))
    with mrc.open(p, "r+") as f:
        f.data[0] = 1
        f.update_header_stats()
    with mrc.open(p) as f:
        assert float(f.header["dmax"]) == 7 and float(f.header["dmin"]) == 0


def test_rejects_bad_files(tmp_path):
    p = tmp_path / "bad.mrc"
    p.write_bytes(b"\0" * 100)
    with pytest.raises(ValueError, match="shorter"):
        mrc.open(p)
    raw = spec_header(8, 8, 2, 2, 8, 8, 2, (8, 8, 2), 1, 0)
    p.write_bytes(raw + b"\0" * 100)
    with pytest.raises(ValueError, match="header describes"):
        mrc.open(p)
    raw = spec_header(8, 8, 1, 3, 8, 8, 1, (8, 8, 1), 0, 0)
    p.write_bytes(raw + b"\0" * 256)
    with pytest.raises(ValueError, match="mode 3"):
        mrc.open(p)


# ---- command line -----------------------------------------------------------

def test_reconstruct_command_options(monkeypatch):
    from guanaco_amd import command_line as cl
    seen = {}
    monkeypatch.setattr(cl, "reconstruct_file", lambda **kw: seen.update(kw))
    cl.main(["-i", "in.mrc", "-o", "out.mrc", "-d", "gpu", "--df", "20000",
             "--ndf", "4", "--Cs", "2.7", "--phi", "5.7", "--angular_weights",
             "False"])
    assert seen["input_filename"] == "in.mrc"
    assert seen["output_filename"] == "out.mrc" and seen["device"] == "gpu"
    assert seen["defocus"] == 20000.0 and seen["num_defocus"] == 4
    assert seen["spherical_aberration"] == 2.7 and seen["phase_shift"] == 5.7
    assert seen["energy"] == 300 and seen["astigmatism"] == 0
    assert seen["corrected_filename"] == "GUANACO_CORRECTED.dat"
    assert seen["method"] == "FBP_CTF" and seen["num_iter"] == 50
    # type=bool, as the reference: the string "False" is truthy
    assert seen["angular_weights"] is True
    assert seen["fused"] is True and seen["storage"] == "f32"
    seen.clear()
    cl.main(["-i", "a", "--device", "gpu", "--storage", "f16", "--depth",
             "64", "--no-fused", "-n", "2"])
    assert seen["device"] == "gpu" and seen["storage"] == "f16"
    assert seen["depth"] == 64 and seen["fused"] is False
    assert seen["ncore"] == 2 and seen["defocus"] is None
    seen.clear()
    cl.main(["-i", "a", "-d,--device", "gpu"])   # the reference's spelling
    assert seen["device"] == "gpu"
    with pytest.raises(SystemExit):
        cl.main(["-o", "x"])   # -i is required


def test_correct_command_options(monkeypatch):
    from guanaco_amd import command_line as cl
    seen = {}
    monkeypatch.setattr(cl, "correct_file", lambda **kw: seen.update(kw))
    cl.correct(["-i", "in.mrc", "-d", "gpu", "--df", "1e4", "--sdf", "200",
                "--ndf", "3", "--centre", "31.5"])
    assert seen["output_filename"] == "corrected.mrc"
    assert seen["step_defocus"] == 200 and seen["num_defocus"] == 3
    assert seen["centre"] == 31.5 and seen["device"] == "gpu"


def test_file_drivers_refuse_cpu_and_other_methods(tmp_path):
    import guanaco_amd as G
    p = tmp_path / "t.mrc"
    out = tmp_path / "o.mrc"
    mrc.write(p, numpy.zeros((3, 4, 8), numpy.float32), voxel_size=1.0)
    with pytest.raises(RuntimeError, match="GPU device"):
        G.reconstruct_file(str(p), str(out), start_angle=-10, step_angle=10,
                           pixel_size=1.0)
    with pytest.raises(RuntimeError, match="GPU device"):
        G.correct_file(str(p), str(out), defocus=1e4)
    with pytest.raises(RuntimeError, match="requires tomopy"):
        G.reconstruct_file(str(p), str(out), start_angle=-10, step_angle=10,
                           pixel_size=1.0, method="SIRT", device="gpu")
    with pytest.raises(ValueError, match="extended header"):
        G.reconstruct_file(str(p), str(out), device="gpu")
    assert not out.exists()


def test_projection_metadata_of_a_corrected_stack(tmp_path):
    from guanaco_amd.detail.reconstruct import read_projection_metadata
    T, ndf = 3, 4
    ext = numpy.zeros((T, ndf), mrc.FEI1_EXTENDED_HEADER_DTYPE)
    ext["Metadata size"] = 768
    ext["Alpha tilt"] = numpy.array([-45.0, 0.0, 45.0])[:, None]
    ext["Defocus"] = numpy.array([-300.0, -100.0, 100.0, 300.0])[None, :]
    p = tmp_path / "c.mrc"
    mrc.write(p, numpy.zeros((T, ndf, 2, 8), numpy.float32), voxel_size=1.5,
              extended_header=ext, exttyp=b"FEI1")
    with mrc.open(p) as f:
        angles, vs, mn, mx = read_projection_metadata(f)
    assert angles.dtype == numpy.float32
    numpy.testing.assert_array_equal(
        angles, numpy.deg2rad([-45.0, 0.0, 45.0]).astype(numpy.float32))
    assert (mn, mx) == (-300.0, 300.0) and float(vs["x"]) == 1.5
