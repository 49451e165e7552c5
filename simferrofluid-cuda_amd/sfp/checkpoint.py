"""The reference's checkpoint file ``<frame>.sav`` (demo/Driver.cpp:95-100,
108-118): Simulation::Save (core/Simulation.cpp:118-125) followed by
StopWatch::Save (core/StopWatch.cpp:18-26). Little-endian, no header:

* velocity face arrays, axis 0, 1, 2: FP64, (n + e_a) values each, z-fastest
* level set: FP64, n^3 values
* initial volume, cumulative volume error, volume error: FP64 x 3
* maximum mesh size: int32
* stopwatch count c: uint64 (size_t), then c NUL-terminated names,
  c call counts (uint32), c mean times (FP64), c maximum times (FP64)

The grid is not stored; readers take the resolution. Arrays come back in the
library's layout (numpy C order on (nx, ny, nz) for cells, (n + e_a) for face
grid a), ready for the solver entry points.
"""
from __future__ import annotations

import struct

import numpy as np


def _face_shape(n, a):
    s = list(n)
    s[a] += 1
    return tuple(s)


def to_bytes(n, velocity, phi, init_volume=0.0, cumul_vol_error=0.0, vol_error=0.0,
             max_mesh_size=0, stopwatch=None) -> bytes:
    n = tuple(int(v) for v in n)
    parts = []
    for a in range(3):
        v = np.ascontiguousarray(velocity[a], dtype="<f8")
        if v.shape != _face_shape(n, a):
            raise ValueError(f"velocity[{a}] has shape {v.shape}, expected {_face_shape(n, a)}")
        parts.append(v.tobytes())
    p = np.ascontiguousarray(phi, dtype="<f8")
    if p.shape != n:
        raise ValueError(f"level set has shape {p.shape}, expected {n}")
    parts.append(p.tobytes())
    parts.append(struct.pack("<dddi", init_volume, cumul_vol_error, vol_error, max_mesh_size))
    sw = stopwatch or {"names": [], "counts": [], "avg": [], "max": []}
    parts.append(struct.pack("<Q", len(sw["names"])))
    for name in sw["names"]:
        parts.append(name.encode() + b"\0")
    parts.append(np.asarray(sw["counts"], dtype="<u4").tobytes())
    parts.append(np.asarray(sw["avg"], dtype="<f8").tobytes())
    parts.append(np.asarray(sw["max"], dtype="<f8").tobytes())
    return b"".join(parts)


def write(path, n, velocity, phi, **kw):
    with open(path, "wb") as f:
        f.write(to_bytes(n, velocity, phi, **kw))


def from_bytes(buf: bytes, n) -> dict:
    n = tuple(int(v) for v in n)
    off = 0

    def take(count, dtype):
        nonlocal off
        dt = np.dtype(dtype)
        size = count * dt.itemsize
        if off + size > len(buf):
            raise ValueError("checkpoint ends early (wrong resolution?)")
        a = np.frombuffer(buf, dtype=dt, count=count, offset=off).copy()
        off += size
        return a

    vel = [take(int(np.prod(_face_shape(n, a))), "<f8").reshape(_face_shape(n, a)) for a in range(3)]
    phi = take(int(np.prod(n)), "<f8").reshape(n)
    init_volume, cumul, vol_error = (float(v) for v in take(3, "<f8"))
    max_mesh = int(take(1, "<i4")[0])
    count = int(take(1, "<u8")[0])
    names = []
    for _ in range(count):
        end = buf.index(b"\0", off)
        names.append(buf[off:end].decode())
        off = end + 1
    counts = take(count, "<u4")
    avg = take(count, "<f8")
    mx = take(count, "<f8")
    return {"velocity": vel, "phi": phi, "init_volume": init_volume, "cumul_vol_error": cumul,
            "vol_error": vol_error, "max_mesh_size": max_mesh,
            "stopwatch": {"names": names, "counts": counts, "avg": avg, "max": mx},
            "trailing_bytes": len(buf) - off}


def read(path, n) -> dict:
    with open(path, "rb") as f:
        return from_bytes(f.read(), n)
