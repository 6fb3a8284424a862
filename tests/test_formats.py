"""MRC and STAR I/O (bsoft_amd/formats.py): round trips and the reference's
rules for stacks, origins and CTF fields (no GPU needed)."""
import math
import os

import numpy as np
import pytest

from bsoft_amd import formats


def test_mrc_round_trip_volume(tmp_path):
    rng = np.random.default_rng(0)
    v = rng.normal(size=(8, 6, 5)).astype(np.float32)
    p = str(tmp_path / "v.mrc")
    formats.write_mrc(p, v, (1.5, 1.5, 1.5), (4, 3, 2))
    m = formats.read_mrc(p)
    assert m.images == 1 and m.data.shape == (1, 8, 6, 5)
    assert np.array_equal(m.data[0], v)
    assert m.sampling == pytest.approx((1.5, 1.5, 1.5))
    assert m.origin == pytest.approx((4, 3, 2))
    with open(p, "rb") as f:
        raw = f.read(1024)
    assert raw[208:212] == b"MAP " and len(raw) == 1024
    assert os.path.getsize(p) == 1024 + v.nbytes


def test_mrc_stack_and_select(tmp_path):
    rng = np.random.default_rng(1)
    st = rng.normal(size=(7, 16, 16)).astype(np.float32)
    p = str(tmp_path / "s.mrcs")
    formats.write_mrc(p, st, (2.0, 2.0, 1.0), (8, 8, 0), images=7)
    m = formats.read_mrc(p)
    assert m.images == 7 and m.data.shape == (7, 1, 16, 16)
    assert np.array_equal(m.data[:, 0], st)
    one = formats.read_mrc(p, img_select=3)
    assert np.array_equal(one.data[0, 0], st[3])
    # pre-2014 header: a .mrcs extension makes a stack (readMRC rules)
    with open(p, "r+b") as f:
        f.seek(108)
        f.write(np.int32(0).tobytes())
    assert formats.read_mrc(p).images == 7


def test_mrc_big_endian_and_modes(tmp_path):
    h = np.zeros((), dtype=formats.MRC_HEADER.newbyteorder(">"))
    h["nx"], h["ny"], h["nz"], h["mode"] = 4, 3, 1, 1
    h["mx"], h["my"], h["mz"] = 4, 3, 1
    h["cella"] = (4.0, 3.0, 1.0)
    h["mapc"], h["mapr"], h["maps"] = 1, 2, 3
    d = np.arange(12, dtype=">i2").reshape(1, 3, 4)
    p = str(tmp_path / "be.mrc")
    with open(p, "wb") as f:
        f.write(h.tobytes())
        f.write(d.tobytes())
    m = formats.read_mrc(p)
    assert np.array_equal(m.data[0, 0], np.arange(12, dtype=np.float32).reshape(3, 4))


CTF_TAGS = {"ctf.voltage": "300000", "ctf.Cs": "2.7e7", "ctf.amp_contrast": "0.07",
            "ctf.defocus_average": "20000", "ctf.defocus_deviation": "500",
            "ctf.astigmatism_angle": "30"}


def _project(tmp_path, n=5, ctf_tags=None):
    rng = np.random.default_rng(2)
    imgs = rng.normal(size=(n, 16, 16)).astype(np.float32)
    formats.write_mrc(str(tmp_path / "parts.mrcs"), imgs, (1.5, 1.5, 1.0), (8, 8, 0), images=n)
    rows = []
    for i in range(n):
        rows.append([i + 1, 1.1e4 + 100 * i, 8.0, 7.5, 0.0, 0.6, 0.8, 30.0 * i, 1 if i != 2 else 0])
    b = formats.StarBlock(name="mg1", items={
        "micrograph.id": "mg1", "micrograph.pixel_size_x": "1.5",
        "particle.file_name": "parts.mrcs", **(CTF_TAGS if ctf_tags is None else ctf_tags)},
        loops=[(["particle.id", "particle.defocus", "particle.origin_x", "particle.origin_y",
                 "particle.view_x", "particle.view_y", "particle.view_z", "particle.view_angle",
                 "particle.select"], rows)])
    p = str(tmp_path / "proj.star")
    formats.write_star(p, [b])
    return p, imgs


def test_star_project_particles(tmp_path):
    p, imgs = _project(tmp_path)
    blocks = formats.read_star(p)
    assert len(blocks) == 1 and blocks[0].items["ctf.voltage"] == "300000"
    parts, srcs = formats.project_particles(p)
    assert len(parts) == 5
    assert parts["def_avg"][1] == pytest.approx(1.11e4)           # the particle's defocus
    assert parts["def_dev"][0] == pytest.approx(500.0)            # the micrograph's astigmatism
    assert parts["ast_angle"][0] == pytest.approx(math.radians(30))
    assert parts["amp_shift"][0] == pytest.approx(math.asin(0.07))
    assert parts["volt"][0] == 3e5 and parts["cs"][0] == 2.7e7
    assert parts["view"][3, 3] == pytest.approx(math.radians(90))
    assert tuple(parts["ori"][0]) == (8.0, 7.5)
    assert parts["sel"][2] == 0
    assert [s.index for s in srcs] == [0, 1, 2, 3, 4]
    got, hdr = formats.load_particle_images(srcs)
    assert np.array_equal(got, imgs)
    assert np.allclose(hdr, 8.0)


def test_star_deprecated_ctf_tags(tmp_path):
    """Old-style projects keep the micrograph CTF under micrograph.voltage and
    micrograph.ctf.* (include/mg_tags.h:129-142); they read as the ctf.* tags
    (rwmgSTAR.cpp:113-126)."""
    old = {"micrograph.voltage": "200000", "micrograph.ctf.Cs": "2.0e7",
           "micrograph.ctf.amp_contrast": "0.1", "micrograph.ctf.defocus_average": "15000",
           "micrograph.ctf.defocus_deviation": "300", "micrograph.ctf.astigmatism_angle": "45"}
    p, _ = _project(tmp_path, ctf_tags=old)
    parts, _ = formats.project_particles(p)
    assert parts["volt"][0] == 2e5 and parts["cs"][0] == 2.0e7
    assert parts["amp_shift"][0] == pytest.approx(math.asin(0.1))
    assert parts["def_dev"][0] == pytest.approx(300.0)
    assert parts["ast_angle"][0] == pytest.approx(math.radians(45))
    assert parts["def_avg"][0] == pytest.approx(1.1e4)         # the particle's own defocus
    # defocus minimum / maximum instead of average / deviation
    p, _ = _project(tmp_path, ctf_tags={"micrograph.voltage": "300000",
                                        "micrograph.ctf.defocus_min": "10000",
                                        "micrograph.ctf.defocus_max": "12000"})
    parts, _ = formats.project_particles(p)
    assert parts["def_dev"][0] == pytest.approx(1000.0)
    assert parts["cs"][0] == 0.0 and parts["amp_shift"][0] == 0.0


def test_star_without_ctf(tmp_path):
    """No voltage tag: the micrograph has no CTF (ctf_from_starblock returns
    NULL) and the default CTFparam applies: 300 kV, no aberrations."""
    p, _ = _project(tmp_path, ctf_tags={"ctf.Cs": "2.7e7"})
    parts, _ = formats.project_particles(p)
    assert parts["volt"][0] == 3e5 and parts["cs"][0] == 0.0 and parts["def_dev"][0] == 0.0
    assert parts["amp_shift"][0] == 0.0


def test_star_syntax(tmp_path):
    p = str(tmp_path / "x.star")
    with open(p, "w") as f:
        f.write("# comment\ndata_a\n_x.one 1\n_x.two 'two words'\nloop_\n_y.a\n_y.b\n"
                "1 2 # trailing\n3 4\n\ndata_b\n_z.t\n;\nmulti\nline\n;\n")
    b = formats.read_star(p)
    assert [x.name for x in b] == ["a", "b"]
    assert b[0].items == {"x.one": "1", "x.two": "two words"}
    assert b[0].loops == [(["y.a", "y.b"], [["1", "2"], ["3", "4"]])]
    assert b[1].items["z.t"] == "multi\nline"
