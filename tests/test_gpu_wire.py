"""GPU: the particle and meshTriangle blocks of the mmap wire format packed on
the device (lf_wire_particles, lf_wire_triangles) against the same blocks
packed on the host from lf_state_download / lf_mc_download, byte for byte,
including into a shared file mapping."""
import os
import tempfile

import numpy as np
import pytest

import libfluid_amd as lf
from libfluid_amd import scenes, wire

pytestmark = pytest.mark.gpu

TS = 1_700_000_000_123


def _host_particles(ctx):
    d = ctx.download()
    recs = wire.particles(d["position"], d["velocity"], d["mass"], d["type"], d["colour"],
                          d["order"])
    return wire.pack_block("particle", recs, TS)


def _host_triangles(tris):
    return wire.pack_block("meshTriangle", wire.mesh_triangles(tris), TS)


@pytest.mark.parametrize("attributes", [0, 1])
def test_wire_blocks_match_host_packing(attributes):
    sc = scenes.dam_break_8k()
    p = sc.params(flags=lf.LF_FLAG_DIFFUSE)
    with lf.Context(0) as ctx:
        ctx.upload(sc.position, sc.velocity, sc.mass, sc.type, sc.colour)
        ctx.step(p, 2)
        dev = ctx.wire_particles(timestamp=TS)
        assert len(dev) == 24 + 44 * len(sc.mass)
        assert dev == _host_particles(ctx)
        if attributes:
            m = lf.mc_reference_params(isolevel=100.0, resolution=1.0, attributes=1)
        else:
            m = sc.mc()
        tris = ctx.mesh(p, m)
        assert len(tris) > 100
        devt = ctx.wire_triangles(timestamp=TS)
        assert devt == _host_triangles(tris)
        h, recs, _ = wire.read_block("meshTriangle", devt)
        assert h["written"] == h["entries"] == len(tris)
        if attributes:
            assert (recs["c.x"] != 0xFFFFFFFF).any()
        else:
            assert (recs["c.x"] == 0xFFFFFFFF).all() and (recs["n0.x"] == 0).all()


def test_wire_into_mapped_file_and_capacity():
    sc = scenes.small_block(n_side=7)
    p = sc.params()
    with lf.Context(0) as ctx:
        ctx.upload(sc.position, sc.velocity, sc.mass, sc.type, sc.colour)
        ctx.step(p)
        need = 24 + 44 * len(sc.mass)
        with tempfile.TemporaryDirectory() as d:
            path = os.path.join(d, "particles.mmf")
            with wire.MappedFile(path, need + 100) as f:
                assert ctx.wire_particles(f.map, timestamp=TS) == need
                f.map.flush()
            with open(path, "rb") as fh:
                data = fh.read()
        assert data[:need] == _host_particles(ctx)
        assert data[need:need + 100] == b" " * 100  # nothing written past the block
        small = bytearray(b"x" * (need - 1))
        with pytest.raises(lf.LfError):
            ctx.wire_particles(small, timestamp=TS)
        assert small == bytearray(b"x" * (need - 1))  # refused before any byte is written
        # empty state: a header with zero entries
        ctx.upload(np.zeros((0, 3), np.float32), np.zeros((0, 3), np.float32),
                   np.zeros(0, np.float32))
        empty = ctx.wire_particles(timestamp=TS)
        h, recs, _ = wire.read_block("particle", empty)
        assert len(empty) == 24 and h["entries"] == h["written"] == 0 and len(recs) == 0
