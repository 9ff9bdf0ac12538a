"""Outputs of the reference's step hook on the host: lattice dumps in
LatticeDump's format (profiles/io/lattice_dump.cpp:14-63), reading them back
as the XYZ reader does (tools/xyz_reader.cpp:15-80, tools/particle.h), dump
file names (fmt::format(dump_file_path, step)) and the `output` section of
config.yaml (config_parsing.cpp:216-229)."""
import os

import numpy as np
import pytest

import akmc


def _readout_bits(s):
    """The flag byte a dump preserves: the readout species only."""
    s = np.asarray(s, dtype=np.uint8)
    out = np.zeros_like(s)
    out[(s & 8) != 0] = 8
    out[(s & 4) != 0] = 4
    out[(s & 2) != 0] = 2
    out[(s & 1) != 0] = 1
    return out


def test_dump_of_one_cell_literal(tmp_path):
    # One unit cell: the corner site (x = 0) and the body site (x = 1), ghost
    # width 2: coordinates of the ghost-extended grid, body site at +0.5.
    p = tmp_path / "one.out"
    akmc.write_dump(np.array([2, 1 | 8], dtype=np.uint8), (1, 1, 1), 2, p)
    assert p.read_text() == "2\n\nCu\t2\t2\t2\nV\t2.5\t2.5\t2.5\n"


def test_dump_format_and_round_trip(tmp_path):
    rng = np.random.default_rng(3)
    box = (5, 4, 3)
    n = 2 * 5 * 4 * 3
    s = rng.choice(np.array([0, 0, 0, 1, 2, 4, 8, 1 | 4, 1 | 8, 2 | 8], dtype=np.uint8), size=n)
    p = tmp_path / "kmc.0.out"
    akmc.write_dump(s, box, 2, p)
    lines = p.read_text().split("\n")
    assert lines[0] == str(n) and lines[1] == "" and lines[-1] == "" and len(lines) == n + 3
    names = {0: "Fe", 1: "V", 2: "Cu", 4: "Mn", 8: "Ni"}
    rb = _readout_bits(s)
    seen = set()
    for i, ln in enumerate(lines[2:-1]):
        t, x, y, z = ln.split("\t")
        X, Y = 10, 4
        xr, yr, zr = i % X, (i // X) % Y, i // (X * Y)   # z, y, x order
        assert t == names[int(rb[i])]
        if xr % 2 == 0:
            assert (x, y, z) == (str(xr // 2 + 2), str(yr + 2), str(zr + 2))
        else:
            assert (float(x), float(y), float(z)) == (xr // 2 + 2.5, yr + 2.5, zr + 2.5)
        seen.add((x, y, z))
    assert len(seen) == n
    r, b, g = akmc.read_dump(p)
    assert b == box and g == 2
    assert np.array_equal(r, rb)
    # reading and writing again reproduces the file
    q = tmp_path / "again.out"
    akmc.write_dump(r, b, g, q)
    assert q.read_text() == p.read_text()
    # the isolated-Cu statistic survives the dump
    assert akmc.cu_clusters(r, b)["isolated"] == akmc.cu_clusters(s, box)["isolated"]


def test_dump_large_coordinates(tmp_path):
    # Body-site coordinates print with six significant digits (ostream default).
    box = (12, 9, 9)
    s = np.zeros(2 * 12 * 9 * 9, dtype=np.uint8)
    p = tmp_path / "g.out"
    akmc.write_dump(s, box, 997, p)
    body = p.read_text().split("\n")[3]
    assert body == "Fe\t997.5\t997.5\t997.5"
    r, b, g = akmc.read_dump(p)
    assert b == box and g == 997


def test_read_dump_errors(tmp_path):
    p = tmp_path / "bad.out"
    p.write_text("3\n\nFe\t2\t2\t2\nV\t2.5\t2.5\t2.5\n")
    with pytest.raises(akmc.AkmcError):
        akmc.read_dump(p)          # count mismatch
    p.write_text("2\n\nFeCu\t2\t2\t2\nV\t2.5\t2.5\t2.5\n")
    with pytest.raises(akmc.AkmcError):
        akmc.read_dump(p)          # dumbbell types have no place in the engine's site bytes
    with pytest.raises(akmc.AkmcError):
        akmc.read_dump(tmp_path / "missing.out")


def test_dump_path():
    assert akmc.dump_path("kmc.{}.out", 100) == "kmc.100.out"
    assert akmc.dump_path("run/{}/lat.{}", 7) == "run/7/lat.7"
    assert akmc.dump_path("a{{b}}.{}", 3) == "a{b}.3"
    with pytest.raises(akmc.AkmcError):
        akmc.dump_path("kmc.{:05d}.out", 1)


def test_output_section(tmp_path):
    p = tmp_path / "config.yaml"
    p.write_text("box:\n  size: [8, 8, 8]\n  lattice_const: 2.85532\n  cutoff_radius: 5.6\n"
                 "output:\n  dump:\n    interval: 100 # dump system lattice information.\n"
                 "    file_path: \"kmc.{}.out\"\n  logs:\n    interval: 10\n    logs_file: \"\"\n")
    o = akmc.output_from_yaml(p)
    assert o == {"dump_interval": 100, "logs_interval": 10, "dump_file_path": "kmc.{}.out", "logs_file": ""}
    q = tmp_path / "noout.yaml"
    q.write_text("box:\n  size: [8, 8, 8]\n")
    assert akmc.output_from_yaml(q)["dump_interval"] == 0


def test_checkpoint_load_rejects_other_files(tmp_path):
    """The checkpoint reader checks the file before it creates an engine (no
    device is touched for a file that is not a checkpoint of this build)."""
    p = tmp_path / "x.ckpt"
    p.write_bytes(b"AKMCCKP0" + bytes(64))
    with pytest.raises(akmc.AkmcError):
        akmc.Engine.load(p)
    with pytest.raises(akmc.AkmcError):
        akmc.Engine.load(tmp_path / "missing.ckpt")


def test_read_dump_rejects_partial_unit_cells(tmp_path):
    """Doubled x must start at a corner site and span whole cells: a dump
    whose x span is odd, or that starts at a body site, is refused (it would
    otherwise map a site past the end of the box)."""
    p = tmp_path / "odd.out"
    xs = [0, 0.5, 1, 1.5, 2, 2.5, 3, 4]  # doubled x 0..6 and 8: span 9
    p.write_text("8\n\n" + "".join("Fe\t%g\t%g\t%g\n" % (x, x % 1, x % 1) for x in xs))
    with pytest.raises(akmc.AkmcError):
        akmc.read_dump(p)
    q = tmp_path / "body_first.out"
    q.write_text("2\n\nFe\t0.5\t0.5\t0.5\nFe\t1\t0\t0\n")  # doubled x 1..2
    with pytest.raises(akmc.AkmcError):
        akmc.read_dump(q)


def test_checkpoint_load_checks_the_site_count(tmp_path):
    """A checkpoint header with a valid magic and sizes but a site count the
    box does not imply, or a short file, is an error status, not an
    allocation failure through the C boundary."""
    import ctypes
    prm = akmc.params_from_config(dict(box=(8, 8, 8), alloy=[("Fe", 1)], va_count=1,
                                       seeds=dict(create_types=1, create_vacancy=2, time_inc=3),
                                       temperature=573.0, attempt_freq=6e12, sector_T=3e-7, physics_time=1.0))
    state_size = 136  # sizeof(akmc::PlainState), akmc_plain.h: 12 x 8-byte fields, 4 x int32, 3 x int64
    hdr = b"AKMCCKP1" + np.array([ctypes.sizeof(akmc.Params), state_size], dtype=np.uint32).tobytes()
    body = bytes(prm) + bytes(state_size)
    for n, tail in [(1 << 50, b""), (2 * 8 * 8 * 8, bytes(100)), (2 * 8 * 8 * 8 + 1, bytes(1025))]:
        p = tmp_path / ("c%d.ckpt" % n)
        p.write_bytes(hdr + body + np.int64(n).tobytes() + tail)
        with pytest.raises(akmc.AkmcError) as ei:
            akmc.Engine.load(p)
        assert "checkpoint" in str(ei.value)
