"""The memory-mapped file wire format of libfluid's sample application
(include/fluid/mmf.hpp:10-187, samples/structures.hpp:11-508) as packed
numpy dtypes, for renderers and tools written in Python.

Every record is its fields packed back to back, little-endian, each as wide
as its C++ type. A block is a 24-byte header (timestamp ms, entries, written)
followed by `entries` records; writers store the header with written = 0,
then the records, then the header with written = entries
(structures.hpp:391-425). scene.mmf is a SceneMeta record and four blocks:
wells, sources, drains, queries (structures.hpp:432-476).

The field tables below follow the DECL_MEMBER lists of structures.hpp
(particle :73-85, well :95-102, source :116-128, drain :138-146, triangle
:174-186, vec3 :193-199, query :211-218, queryResult :221-230, meshTriangle
:234-258, header :284-290, sceneMeta :373-388).
"""
import json
import mmap
import os
import time

import numpy as np

_F4, _U4, _I4, _U8, _I8, _B1 = "<f4", "<u4", "<i4", "<u8", "<i8", "u1"


def _xyz(prefix, t=_F4):
    return [(prefix + ".x", t), (prefix + ".y", t), (prefix + ".z", t)]


LAYOUTS = {
    "header": [("timestamp", _I8), ("entries", _U8), ("written", _U8)],
    "well": _xyz("well") + [("force", _F4)],
    "source": _xyz("source") + _xyz("source.vel") + [("rate", _U8), ("tag", _U8),
                                                     ("colour", _U4)],
    "drain": _xyz("drain") + [("width", _F4), ("depth", _F4)],
    "sceneMeta": [("suspend", _B1), ("terminate", _B1), ("solverIter", _I4),
                  ("solverStep", _F4), ("solverScale", _F4), ("surfaceRes", _F4),
                  ("gravity", _F4)] + _xyz("minBound") + _xyz("maxBound"),
    "vec3": _xyz("vec3"),
    "query": [("query.id", _U8)] + _xyz("query"),
    "queryResult": [("query.id", _U8)] + _xyz("query") + [("neighbours", _U8),
                                                          ("avgcolour", _U4)],
    "particle": [("id", _U8), ("type", _I4), ("mass", _F4), ("colour", _U4)] +
                _xyz("position") + _xyz("velocity"),
    "triangle": _xyz("v0") + _xyz("v1") + _xyz("v2"),
    "meshTriangle": (_xyz("v0") + _xyz("v1") + _xyz("v2") + _xyz("n0") + _xyz("n1") +
                     _xyz("n2") + _xyz("c", _U4)),
}

DTYPES = {k: np.dtype(v) for k, v in LAYOUTS.items()}  # packed: numpy adds no padding
HEADER_BYTES = DTYPES["header"].itemsize


def defs():
    """The defs.json object the sample writes (samples/main.cpp:146-182)."""
    return {k: {"fields": [{"name": n, "size": np.dtype(t).itemsize} for n, t in v],
                "size": DTYPES[k].itemsize} for k, v in LAYOUTS.items()}


def defs_json():
    """defs() as nlohmann::json::dump(1) prints it (sorted keys, one-space indent)."""
    return json.dumps(defs(), indent=1, sort_keys=True, separators=(",", ": "))


def now_ms():
    return int(time.time() * 1000)


def pack_block(kind, records, timestamp=None):
    """bytes of one block: header (written = entries) + records."""
    records = np.ascontiguousarray(records, dtype=DTYPES[kind])
    h = np.zeros(1, DTYPES["header"])
    h["timestamp"] = now_ms() if timestamp is None else timestamp
    h["entries"] = h["written"] = len(records)
    return h.tobytes() + records.tobytes()


def read_block(kind, buf, offset=0):
    """(header, records, offset after the block) from a bytes-like object."""
    buf = memoryview(buf).cast("B")
    h = np.frombuffer(buf, DTYPES["header"], 1, offset)[0]
    offset += HEADER_BYTES
    n = int(h["entries"])
    recs = np.frombuffer(buf, DTYPES[kind], n, offset).copy()
    return h.copy(), recs, offset + n * DTYPES[kind].itemsize


def write_block(kind, buf, records, timestamp=None, offset=0):
    """Writes a block into a writable buffer (e.g. an mmap) in the reference's
    order: header with written = 0, the records, then written = entries.
    Returns the offset after the block."""
    records = np.ascontiguousarray(records, dtype=DTYPES[kind])
    mv = memoryview(buf).cast("B")
    end = offset + HEADER_BYTES + records.nbytes
    if end > len(mv):
        raise ValueError("wire: block of %d bytes does not fit at offset %d of %d" % (
            end - offset, offset, len(mv)))
    h = np.zeros(1, DTYPES["header"])
    h["timestamp"] = now_ms() if timestamp is None else timestamp
    h["entries"] = len(records)
    mv[offset:offset + HEADER_BYTES] = h.tobytes()
    mv[offset + HEADER_BYTES:end] = records.tobytes()
    h["written"] = len(records)
    mv[offset:offset + HEADER_BYTES] = h.tobytes()
    return end


def particles(position, velocity, mass, ptype, colour, ids):
    """Particle records from state arrays (as Context.download returns them)."""
    n = len(mass)
    r = np.zeros(n, DTYPES["particle"])
    r["id"] = ids
    r["type"] = ptype
    r["mass"] = mass
    r["colour"] = colour
    for k, a in enumerate("xyz"):
        r["position." + a] = np.asarray(position)[:, k]
        r["velocity." + a] = np.asarray(velocity)[:, k]
    return r


def mesh_triangles(tris):
    """meshTriangle records from a lf_mesh_triangle array (triangle_dtype)."""
    r = np.zeros(len(tris), DTYPES["meshTriangle"])
    for j in range(3):
        for k, a in enumerate("xyz"):
            r["v%d.%s" % (j, a)] = tris["v"][:, j, k]
            r["n%d.%s" % (j, a)] = tris["n"][:, j, k]
        r["c." + "xyz"[j]] = tris["c"][:, j]
    return r


def read_scene(buf):
    """scene.mmf: (meta record, wells, sources, drains, queries)."""
    meta = np.frombuffer(memoryview(buf).cast("B"), DTYPES["sceneMeta"], 1, 0)[0].copy()
    off = DTYPES["sceneMeta"].itemsize
    out = [meta]
    for kind in ("well", "source", "drain", "query"):
        _, recs, off = read_block(kind, buf, off)
        out.append(recs)
    return tuple(out)


def write_scene(buf, meta, wells=(), sources=(), drains=(), queries=(), timestamp=None):
    mv = memoryview(buf).cast("B")
    m = np.ascontiguousarray(meta, DTYPES["sceneMeta"]).reshape(1)
    mv[:m.nbytes] = m.tobytes()
    off = m.nbytes
    for kind, recs in (("well", wells), ("source", sources), ("drain", drains),
                       ("query", queries)):
        recs = np.asarray(recs, DTYPES[kind]) if len(recs) else np.zeros(0, DTYPES[kind])
        off = write_block(kind, buf, recs, timestamp, off)
    return off


class MappedFile:
    """A file padded with max(4096, length) spaces and mapped shared
    (miommf::padOut + createSink, structures.hpp:13-49)."""

    def __init__(self, path, length=None, writable=True):
        if writable:
            size = max(4096, int(length or 0))
            with open(path, "wb") as f:
                f.write(b" " * size)
        self._f = open(path, "r+b" if writable else "rb")
        size = os.fstat(self._f.fileno()).st_size
        self.map = mmap.mmap(self._f.fileno(), size,
                             access=mmap.ACCESS_WRITE if writable else mmap.ACCESS_READ)

    def close(self):
        self.map.close()
        self._f.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
