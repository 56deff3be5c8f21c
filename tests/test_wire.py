"""The mmap wire format (include/fluid/wire.hpp, libfluid_amd/wire.py) against
the reference's own record definitions and across the C++ and Python sides.

The record field lists are pinned against the DECL_MEMBER lists of
samples/structures.hpp and the defs.json keys of samples/main.cpp (read as
text when the reference tree is present); the field widths are pinned by
tests/cpp/wire_main.cpp against sizeof each member of the fluid:: records.
"""
import os
import re
import subprocess
import tempfile

import numpy as np
import pytest

from libfluid_amd import wire

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "libfluid_amd", "bin", "wire_main")
REF = "/root/reference/samples"
needs_ref = pytest.mark.skipif(not os.path.isdir(REF), reason="reference tree absent")


@needs_ref
def test_field_lists_follow_reference_structures():
    src = open(os.path.join(REF, "structures.hpp")).read()
    names = dict(re.findall(r'static const char (\w+)\[\] = "([^"]*)";', src))
    defs = {}
    for m in re.finditer(r"auto (\w+)Def\(\) \{(.*?)\n\t\}", src, re.S):
        defs[m.group(1)] = [names[n] for n in re.findall(r"DECL_MEMBER\((\w+),", m.group(2))]
    main = open(os.path.join(REF, "main.cpp")).read()
    # {"key", xType.second} pairs and xType = writeMetaPacked<decltype(strucures::yDef...
    types = dict(re.findall(r"auto (\w+)Type = writeMetaPacked<decltype\(strucures::(\w+)Def",
                            main))
    keys = dict(re.findall(r'\{"(\w+)",\s+(\w+)Type\.second\}', main))
    assert set(keys) == set(wire.LAYOUTS)
    for key, tname in keys.items():
        ref_fields = defs[types[tname]]
        ours = [n for n, _ in wire.LAYOUTS[key]]
        assert ours == ref_fields, key


def test_cpp_widths_and_defs_json():
    with tempfile.TemporaryDirectory() as d:
        r = subprocess.run([EXE, "write", d], capture_output=True, text=True)
        assert r.returncode == 0, r.stderr
        assert "roundtrip ok" in r.stdout
        assert open(os.path.join(d, "defs.json")).read() == wire.defs_json()
    sizes = {k: v.itemsize for k, v in wire.DTYPES.items()}
    assert sizes == {"header": 24, "well": 16, "source": 44, "drain": 20, "sceneMeta": 46,
                     "vec3": 12, "query": 20, "queryResult": 32, "particle": 44,
                     "triangle": 36, "meshTriangle": 84}


def test_cpp_writer_python_reader():
    with tempfile.TemporaryDirectory() as d:
        subprocess.run([EXE, "write", d], check=True, capture_output=True)
        with wire.MappedFile(os.path.join(d, "particles.mmf"), writable=False) as f:
            h, p, end = wire.read_block("particle", f.map)
            assert os.path.getsize(os.path.join(d, "particles.mmf")) == 4096  # padded
            assert bytes(f.map[end:end + 4]) == b"    "  # the padding spaces follow
        assert h["entries"] == h["written"] == 5 and h["timestamp"] > 1_600_000_000_000
        i = np.arange(5)
        assert (p["id"] == 1000 + i).all() and (p["type"] == i % 2).all()
        assert (p["mass"] == 1 + i).all() and (p["colour"] == 0xA0B0C0D0 + i).all()
        assert (p["position.y"] == 2.0 * i).all() and (p["velocity.z"] == -0.125).all()
        with wire.MappedFile(os.path.join(d, "triangles.mmf"), writable=False) as f:
            h, t, _ = wire.read_block("meshTriangle", f.map)
        assert len(t) == 3 and (t["v1.y"] == [0, 1, 2]).all() and (t["n2.z"] == 1).all()
        assert (t["c.x"] == [0x11111111, 0x22222222, 0x33333333]).all()
        with wire.MappedFile(os.path.join(d, "query.mmf"), writable=False) as f:
            _, q, _ = wire.read_block("queryResult", f.map)
        assert q["query.id"].tolist() == [11, 13] and q["neighbours"].tolist() == [27, 5]
        assert q["avgcolour"].tolist() == [0x80402010, 0xFFFFFFFF]
        with wire.MappedFile(os.path.join(d, "scene.mmf"), writable=False) as f:
            meta, wells, sources, drains, queries = wire.read_scene(f.map)
        assert meta["terminate"] == 1 and meta["solverIter"] == 5
        assert np.float32(meta["gravity"]) == np.float32(9.8) and meta["maxBound.z"] == 800
        assert wells["force"].tolist() == [4.5, -4.5]
        assert sources["tag"][0] == 0x123456789 and sources["rate"][0] == 18
        assert drains["width"][0] == 0.75 and queries["query.id"].tolist() == [11, 12, 13]
        with wire.MappedFile(os.path.join(d, "colliders.mmf"), writable=False) as f:
            _, tri, _ = wire.read_block("triangle", f.map)
        assert tri["v1.x"].tolist() == [1.0]
        with wire.MappedFile(os.path.join(d, "dynamic_bodies.mmf"), writable=False) as f:
            _, v, _ = wire.read_block("vec3", f.map)
        assert v["vec3.z"].tolist() == [3.0, 6.0]


def test_python_writer_cpp_reader():
    meta = np.zeros(1, wire.DTYPES["sceneMeta"])
    meta["suspend"], meta["solverIter"], meta["solverStep"] = 1, 7, 0.25
    meta["solverScale"], meta["surfaceRes"], meta["gravity"] = 500, 2, 9.8
    meta["minBound.x"], meta["maxBound.y"] = -3, 42
    wells = np.zeros(1, wire.DTYPES["well"])
    wells["well.y"], wells["force"] = 5, 1.25
    src = np.zeros(2, wire.DTYPES["source"])
    src["source.x"], src["source.vel.z"] = [1, 2], [-1, -2]
    src["rate"], src["tag"], src["colour"] = [9, 16], [2 ** 40, 3], [0xFF0000FF, 1]
    drains = np.zeros(0, wire.DTYPES["drain"])
    queries = np.zeros(2, wire.DTYPES["query"])
    queries["query.id"], queries["query.z"] = [5, 6], [0.5, 0.75]
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "scene.mmf")
        with wire.MappedFile(path, 1 << 14) as f:
            wire.write_scene(f.map, meta, wells, src, drains, queries)
        out = subprocess.run([EXE, "read", path], check=True, capture_output=True,
                             text=True).stdout.splitlines()
    assert out[0].split() == ["meta", "1", "0", "7", "0.25", "500", "2",
                              "%.9g" % np.float32(9.8), "-3", "0", "0", "0", "42", "0"]
    assert out[1].split() == ["well", "0", "5", "0", "1.25"]
    assert out[2].split() == ["source", "1", "0", "0", "0", "0", "-1", "9",
                              str(2 ** 40), str(0xFF0000FF)]
    assert out[3].split()[7:] == ["16", "3", "1"]
    assert [l.split()[1] for l in out[4:]] == ["5", "6"] and len(out) == 6


def test_block_header_protocol():
    buf = bytearray(b" " * 256)
    recs = np.zeros(3, wire.DTYPES["vec3"])
    recs["vec3.x"] = [1, 2, 3]
    end = wire.write_block("vec3", buf, recs, timestamp=1234, offset=8)
    assert end == 8 + 24 + 36
    h, r, off = wire.read_block("vec3", buf, 8)
    assert (h["timestamp"], h["entries"], h["written"]) == (1234, 3, 3) and off == end
    assert bytes(buf[:8]) == b" " * 8 and bytes(buf[end:end + 8]) == b" " * 8
    assert wire.pack_block("vec3", recs, 1234) == bytes(buf[8:end])
    with pytest.raises(ValueError):
        wire.write_block("vec3", bytearray(40), recs)
