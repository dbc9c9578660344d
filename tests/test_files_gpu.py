"""The file drivers and commands end to end on the GPU: MRC tilt series with
FEI1 tilt angles in, MRC volume out (reconstruct_file, ``guanaco -d gpu``),
and MRC corrected stacks (correct_file, ``guanaco.correct -d gpu``), against
the oracle chain correct_projections -> reconstruct on the same data."""
import numpy
import pytest

import guanaco_amd as G
from guanaco_amd import command_line, mrc
from oracle import oracle as O

pytestmark = pytest.mark.gpu

PS, DF, CS, PHI = 1.5, 25000.0, 2.7, 5.7392


def rel_l2(a, b):
    a = numpy.asarray(a, numpy.float64)
    b = numpy.asarray(b, numpy.float64)
    return float(numpy.linalg.norm(a - b) / numpy.linalg.norm(b))


def write_tilts(path, T=9, Y=8, X=64, seed=31):
    rng = numpy.random.default_rng(seed)
    proj = (rng.standard_normal((T, Y, X)) + 5.0).astype(numpy.float32)
    tilts = numpy.linspace(-60.0, 60.0, T)
    ext = numpy.zeros(T, mrc.FEI1_EXTENDED_HEADER_DTYPE)
    ext["Metadata size"] = ext.dtype.itemsize
    ext["Alpha tilt"] = tilts
    with mrc.new_mmap(path, proj.shape, 2, extended_header=ext,
                      exttyp=b"FEI1") as f:
        f.data[...] = proj
        f.set_image_stack()
        f.voxel_size = PS
    angles = numpy.deg2rad(tilts).astype(numpy.float32)
    return proj, angles


def oracle_volume(proj, angles, ndf):
    centre = O.get_centre(proj.shape, None, False)
    c, mn, mx = O.correct_projections(proj, centre, PS, 300, DF, ndf, None,
                                      CS, 0, 0, PHI)
    return c, mn, mx, O.reconstruct(c, angles, None, PS, mn, mx)


@pytest.mark.parametrize("fused", [True, False])
def test_reconstruct_file_matches_oracle(capi, tmp_path, fused):
    src, out = tmp_path / "tilts.mrc", tmp_path / "rec.mrc"
    proj, angles = write_tilts(src)
    G.reconstruct_file(str(src), str(out), str(tmp_path / "corr.dat"),
                       energy=300, defocus=DF, num_defocus=3,
                       spherical_aberration=CS, phase_shift=PHI,
                       device="gpu", fused=fused)
    _, _, _, expect = oracle_volume(proj, angles, 3)
    with mrc.open(out) as f:
        assert f.data.shape == (8, 64, 64) and f.data.dtype == numpy.float32
        assert float(f.voxel_size["x"]) == PS
        assert rel_l2(f.data, expect) < 1e-5
    # the two-step path leaves the corrected intermediate, the fused one not
    assert (tmp_path / "corr.dat").exists() == (not fused)


def test_correct_file_then_reconstruct_corrected_stack(capi, tmp_path):
    src, cor, out = (tmp_path / n for n in ("t.mrc", "c.mrc", "r.mrc"))
    proj, angles = write_tilts(src, T=7, Y=8, X=32, seed=32)
    ndf = 4
    G.correct_file(str(src), str(cor), energy=300, defocus=DF,
                   num_defocus=ndf, spherical_aberration=CS, phase_shift=PHI,
                   device="gpu")
    c, mn, mx, expect = oracle_volume(proj, angles, ndf)
    with mrc.open(cor) as f:
        assert f.data.shape == (7, ndf, 8, 32) and f.is_volume_stack()
        assert rel_l2(f.data, c) < 1e-5
        h = f.indexed_extended_header.reshape(7, ndf)
        numpy.testing.assert_allclose(h["Alpha tilt"][:, 0],
                                      numpy.rad2deg(angles), rtol=1e-6)
        step = (mx - mn) / (ndf - 1)
        numpy.testing.assert_allclose(h["Defocus"][0],
                                      mn + numpy.arange(ndf) * step)
        assert (h["Metadata size"] == 768).all()
    G.reconstruct_file(str(cor), str(out), device="gpu")
    with mrc.open(out) as f:
        assert rel_l2(f.data, expect) < 1e-5


def test_commands_end_to_end(capi, tmp_path):
    src, out = tmp_path / "t.mrc", tmp_path / "r.mrc"
    proj, angles = write_tilts(src, seed=33)
    command_line.main(["-i", str(src), "-o", str(out), "-d", "gpu", "--df",
                       str(DF), "--ndf", "2", "--Cs", str(CS), "--phi",
                       str(PHI), "--depth", "16"])
    _, _, _, expect = oracle_volume(proj, angles, 2)
    with mrc.open(out) as f:
        assert f.data.shape == (8, 16, 64)
        assert rel_l2(f.data, expect[:, 24:40]) < 1e-5
    cor = tmp_path / "c.mrc"
    command_line.correct(["-i", str(src), "-o", str(cor), "-d", "gpu",
                          "--df", str(DF), "--Cs", str(CS), "--phi",
                          str(PHI)])
    with mrc.open(cor) as f:
        assert f.data.shape == (9, 1, 8, 64)


def test_angles_from_the_command_line(capi, tmp_path):
    # no extended header: angles from start/step, no CTF (plain FBP)
    src, out = tmp_path / "t.mrc", tmp_path / "r.mrc"
    rng = numpy.random.default_rng(34)
    proj = rng.standard_normal((11, 4, 32)).astype(numpy.float32)
    mrc.write(src, proj, voxel_size=2.0)
    G.reconstruct_file(str(src), str(out), start_angle=-50.0, step_angle=10.0,
                       pixel_size=2.0, device="gpu", angular_weights=True,
                       transform="minus")
    angles = numpy.deg2rad(-50.0 + numpy.arange(11) * 10.0)
    expect = O.reconstruct(proj, angles.astype(numpy.float32), None, 2.0,
                           angular_weights=True, transform="minus")
    with mrc.open(out) as f:
        assert rel_l2(f.data, expect) < 1e-5
