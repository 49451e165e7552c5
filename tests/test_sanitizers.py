"""Host-code sanitizers on the CPU: the FP64 oracle built with
AddressSanitizer and UndefinedBehaviorSanitizer (oracle/_san/liboracle.so)
runs the oracle's own test modules, and the checkpoint reader/writer driver
built the same way writes and reads a file. Any report ends the run with a
nonzero status. (The drop-in driver, which needs the GPU, runs under
UndefinedBehaviorSanitizer in tests/test_gpu_properties.py.)"""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE = os.path.join(ROOT, "oracle")
SAN_LIB = os.path.join(ORACLE, "_san", "liboracle.so")
CPP = os.path.join(ROOT, "tests", "cpp")


def _asan_runtime():
    out = subprocess.run(["gcc", "-print-file-name=libasan.so"], capture_output=True, text=True)
    path = out.stdout.strip()
    if out.returncode != 0 or not os.path.isabs(path) or not os.path.exists(path):
        pytest.skip("no AddressSanitizer runtime for gcc here")
    return path


def _env():
    env = dict(os.environ)
    # the interpreter is not instrumented: the runtime goes first; Python's
    # own allocations at exit are not leaks of the code under test
    env["LD_PRELOAD"] = _asan_runtime()
    env["ASAN_OPTIONS"] = "detect_leaks=0:abort_on_error=1"
    env["UBSAN_OPTIONS"] = "print_stacktrace=1:halt_on_error=1"
    return env


def test_oracle_suite_under_asan_and_ubsan():
    subprocess.run(["make", "-s", "san"], cwd=ORACLE, check=True)
    nm = subprocess.run(["nm", "-D", SAN_LIB], capture_output=True, text=True, check=True).stdout
    assert "__asan_report_load8" in nm and "__ubsan_handle" in nm  # instrumented
    env = _env()
    env["SFP_ORACLE_LIB"] = SAN_LIB
    probe = subprocess.run([sys.executable, "-c", "import oracle; print(oracle.lib()._name)"],
                           cwd=ORACLE, env=env, capture_output=True, text=True)
    assert probe.returncode == 0, probe.stderr
    assert probe.stdout.strip() == SAN_LIB
    mods = ["tests/test_oracle.py", "tests/test_oracle_closed_forms.py", "tests/test_post_oracle.py",
            "tests/test_unet_oracle.py"]
    r = subprocess.run([sys.executable, "-m", "pytest", *mods, "-q", "-x", "-m", "not gpu",
                        "-p", "no:cacheprovider"], cwd=ROOT, env=env, capture_output=True, text=True)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-3000:]
    assert "passed" in r.stdout


def test_checkpoint_driver_under_asan_and_ubsan(tmp_path):
    subprocess.run(["make", "-s", "checkpoint_test_san"], cwd=CPP, check=True)
    exe = os.path.join(CPP, "checkpoint_test_san")
    env = dict(os.environ, ASAN_OPTIONS="abort_on_error=1", UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    path = str(tmp_path / "ck.sav")
    subprocess.run([exe, "write", path], check=True, env=env)
    out = subprocess.run([exe, "read", path], check=True, env=env, capture_output=True, text=True)
    assert out.stdout.split()
