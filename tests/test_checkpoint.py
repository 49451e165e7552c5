"""The reference checkpoint format (demo/Driver.cpp:95-118,
core/Simulation.cpp:118-136, core/StopWatch.cpp:18-43, core/Common.h:70-91):
bytes written by the Python and C++ writers equal a layout assembled field by
field with struct, and each reader reads what the other wrote."""
import os
import struct
import subprocess

import numpy as np
import pytest

from sfp import checkpoint

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CPP = os.path.join(ROOT, "tests", "cpp", "checkpoint_test")
N = (5, 4, 3)


def seeded_state():
    vel = []
    for a in range(3):
        shape = list(N)
        shape[a] += 1
        cnt = int(np.prod(shape))
        vel.append((1000.0 * a + 0.5 * np.arange(cnt)).reshape(shape))
    phi = (-1.0 + 0.125 * np.arange(int(np.prod(N)))).reshape(N)
    sw = {"names": ["Advect", "Project"], "counts": [3, 7], "avg": [0.25, 1.5], "max": [0.5, 2.0]}
    return vel, phi, dict(init_volume=0.3125, cumul_vol_error=-1.5e-3, vol_error=2.5e-4,
                          max_mesh_size=4242, stopwatch=sw)


def expected_bytes(vel, phi, kw):
    """The layout written out field by field: IO::Write of each GridData's
    vector (raw doubles), three doubles and an int, then StopWatch::Save: the
    size_t count, names each followed by std::ends, the three vectors."""
    b = b""
    for a in range(3):
        b += struct.pack(f"<{vel[a].size}d", *vel[a].ravel())
    b += struct.pack(f"<{phi.size}d", *phi.ravel())
    b += struct.pack("<d", kw["init_volume"]) + struct.pack("<d", kw["cumul_vol_error"])
    b += struct.pack("<d", kw["vol_error"]) + struct.pack("<i", kw["max_mesh_size"])
    sw = kw["stopwatch"]
    b += struct.pack("<Q", len(sw["names"]))
    for name in sw["names"]:
        b += name.encode() + b"\x00"
    b += struct.pack(f"<{len(sw['counts'])}I", *sw["counts"])
    b += struct.pack(f"<{len(sw['avg'])}d", *sw["avg"])
    b += struct.pack(f"<{len(sw['max'])}d", *sw["max"])
    return b


def test_python_writer_matches_layout_and_reads_back():
    vel, phi, kw = seeded_state()
    buf = checkpoint.to_bytes(N, vel, phi, **kw)
    assert buf == expected_bytes(vel, phi, kw)
    got = checkpoint.from_bytes(buf, N)
    for a in range(3):
        np.testing.assert_array_equal(got["velocity"][a], vel[a])
    np.testing.assert_array_equal(got["phi"], phi)
    assert got["init_volume"] == 0.3125 and got["max_mesh_size"] == 4242 and got["trailing_bytes"] == 0
    assert got["stopwatch"]["names"] == ["Advect", "Project"]
    np.testing.assert_array_equal(got["stopwatch"]["counts"], [3, 7])


def test_wrong_resolution_is_caught():
    vel, phi, kw = seeded_state()
    buf = checkpoint.to_bytes(N, vel, phi, **kw)
    with pytest.raises(ValueError):
        checkpoint.from_bytes(buf, (9, 9, 9))
    # a smaller resolution parses but leaves bytes over, or fails on the names
    try:
        got = checkpoint.from_bytes(buf, (4, 4, 3))
        assert got["trailing_bytes"] != 0 or got["stopwatch"]["names"] != ["Advect", "Project"]
    except (ValueError, UnicodeDecodeError):
        pass


def _cpp():
    if not os.path.exists(CPP):
        subprocess.run(["make", "-C", os.path.dirname(CPP), "checkpoint_test"], check=True,
                       capture_output=True)
    return CPP


def test_cpp_writer_bytes_equal_python(tmp_path):
    path = str(tmp_path / "0.sav")
    subprocess.run([_cpp(), "write", path], check=True)
    vel, phi, kw = seeded_state()
    with open(path, "rb") as f:
        assert f.read() == expected_bytes(vel, phi, kw)


def test_cpp_reader_reads_python_file(tmp_path):
    vel, phi, kw = seeded_state()
    vel = [v * 3 - 1 for v in vel]
    kw["stopwatch"]["names"] = ["Step"]
    kw["stopwatch"]["counts"] = [11]
    kw["stopwatch"]["avg"] = [0.125]
    kw["stopwatch"]["max"] = [0.75]
    path = str(tmp_path / "1.sav")
    checkpoint.write(path, N, vel, phi, **kw)
    out = subprocess.run([_cpp(), "read", path], check=True, capture_output=True, text=True).stdout.split()
    assert float(out[0]) == sum(float(v.sum()) for v in vel)
    assert float(out[1]) == float(phi.sum())
    assert out[5:] == ["4242", "1", "Step", "11", "0.125", "0.75"]
