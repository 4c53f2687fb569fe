"""3DGS PLY reading and writing (crates/nano-splat/src/ply.rs), host only."""
import struct

import numpy as np
import pytest

from nanotracer_amd import _lib as L
from nanotracer_amd import synth
from nanotracer_amd.api import read_ply, write_ply


def test_write_then_read_roundtrip(tmp_path):
    # ply.rs:262-287
    s = synth.HostSplats.zeros(1)
    s.pos[0] = (1.0, 2.0, 3.0)
    s.sh_dc[0] = (0.4, 0.5, 0.6)
    s.sh_rest[0] = np.arange(45, dtype=np.float32) * np.float32(0.1)
    s.logit[0] = 2.7
    s.log_scale[0] = (-0.5, -0.6, -0.7)
    s.rot[0] = (1.0, 0.0, 0.0, 0.0)
    p = tmp_path / "roundtrip.ply"
    write_ply(p, s)
    b = read_ply(p)
    assert b.n == 1
    for x, y in zip(b.arrays(), s.arrays()):
        assert np.array_equal(x, y)
    assert np.abs(b.sh_rest[0] - np.arange(45) * 0.1).max() < 1e-5


def test_header_matches_the_reference_writer(tmp_path):
    p = tmp_path / "h.ply"
    write_ply(p, synth.make_scene(3, 1))
    head = p.read_bytes().split(b"end_header\n")[0].decode().splitlines()
    props = [l.split()[-1] for l in head if l.startswith("property float ")]
    assert head[:3] == ["ply", "format binary_little_endian 1.0", "element vertex 3"]
    assert props == (["x", "y", "z", "nx", "ny", "nz", "f_dc_0", "f_dc_1", "f_dc_2"]
                     + [f"f_rest_{i}" for i in range(45)]
                     + ["opacity", "scale_0", "scale_1", "scale_2", "rot_0", "rot_1", "rot_2", "rot_3"])
    assert p.stat().st_size == len("\n".join(head)) + len("\nend_header\n") + 3 * 62 * 4


def test_scene_roundtrip_bit_exact(tmp_path):
    s = synth.make_scene(10_000, 42, 3.0)
    p = tmp_path / "scene.ply"
    write_ply(p, s)
    b = read_ply(p)
    for x, y in zip(b.arrays(), s.arrays()):
        assert np.array_equal(x.view(np.uint32), y.view(np.uint32))


def _write_custom(path, names, rows, fmt="binary_little_endian 1.0", extra=""):
    hdr = f"ply\nformat {fmt}\ncomment made by a test\nelement vertex {len(rows)}\n{extra}"
    hdr += "".join(f"property float {n}\n" for n in names) + "end_header\n"
    with open(path, "wb") as f:
        f.write(hdr.encode())
        for r in rows:
            f.write(struct.pack("<%df" % len(r), *r))


def test_properties_are_mapped_by_name(tmp_path):
    # shuffled order, an unknown property, no normals, only some f_rest
    names = ["rot_3", "opacity", "z", "junk", "f_dc_2", "y", "scale_1", "rot_0", "f_rest_7", "x", "scale_0",
             "f_dc_0", "rot_2", "f_dc_1", "scale_2", "rot_1", "f_rest_44"]
    vals = {n: float(k + 1) for k, n in enumerate(names)}
    p = tmp_path / "shuffled.ply"
    _write_custom(p, names, [[vals[n] for n in names]])
    b = read_ply(p)
    assert tuple(b.pos[0]) == (vals["x"], vals["y"], vals["z"])
    assert tuple(b.rot[0]) == (vals["rot_0"], vals["rot_1"], vals["rot_2"], vals["rot_3"])
    assert tuple(b.log_scale[0]) == (vals["scale_0"], vals["scale_1"], vals["scale_2"])
    assert b.logit[0] == vals["opacity"] and tuple(b.sh_dc[0]) == (vals["f_dc_0"], vals["f_dc_1"], vals["f_dc_2"])
    expect = np.zeros(45, np.float32)
    expect[7], expect[44] = vals["f_rest_7"], vals["f_rest_44"]
    assert np.array_equal(b.sh_rest[0], expect)


@pytest.mark.parametrize("case", ["magic", "ascii", "missing", "list"])
def test_malformed_files_are_rejected(tmp_path, case):
    p = tmp_path / f"{case}.ply"
    full = ["x", "y", "z", "f_dc_0", "f_dc_1", "f_dc_2", "opacity", "scale_0", "scale_1", "scale_2",
            "rot_0", "rot_1", "rot_2", "rot_3"]
    if case == "magic":
        p.write_bytes(b"plx\nformat binary_little_endian 1.0\nend_header\n")
    elif case == "ascii":
        _write_custom(p, full, [[0.0] * len(full)], fmt="ascii 1.0")
    elif case == "missing":
        _write_custom(p, full[:-1], [[0.0] * (len(full) - 1)])
    else:
        _write_custom(p, full, [[0.0] * len(full)], extra="property list uchar int vertex_indices\n")
    with pytest.raises(L.NtsError):
        read_ply(p)
