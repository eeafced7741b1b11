"""The host-only parts of the library (geometry encoder and decoder, DEN/NRRD file I/O, phantom
generator, the serving arena's allocator) and the CPU oracle, built with AddressSanitizer and
UndefinedBehaviorSanitizer, and the threaded parts (geometry from four threads, the oracle's
worker threads) with ThreadSanitizer, then run by tests/cpp/host_sanitize.cpp. No device code is
built here: these runs need no GPU."""
import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
CSRC = ROOT / "ctl_amd" / "csrc"
HOST = [CSRC / "geometry.cpp", CSRC / "fileio.cpp", CSRC / "phantom.cpp", CSRC / "errors.cpp"]
ORACLE = [ROOT / "oracle" / "ctl_oracle.c", ROOT / "oracle" / "ctl_oracle_processing.c"]


def build(tmp: Path, san: str) -> Path:
    if not shutil.which("g++") or not shutil.which("gcc"):
        pytest.skip("no GNU compilers")
    flags = [f"-fsanitize={san}", "-fno-omit-frame-pointer", "-g", "-O1"]
    if san != "thread":
        flags.append("-fno-sanitize-recover=all")
    objs = []
    for c in ORACLE:
        o = tmp / (c.stem + ".o")
        subprocess.run(["gcc", "-std=c11", "-ffp-contract=off", *flags, "-c", str(c), "-o", str(o)], check=True)
        objs.append(str(o))
    exe = tmp / f"host_sanitize_{san.split(',')[0]}"
    cmd = ["g++", "-std=c++17", *flags, f"-I{ROOT / 'include'}", f"-I{CSRC}", f"-I{ROOT / 'oracle'}",
           str(ROOT / "tests" / "cpp" / "host_sanitize.cpp"), *map(str, HOST), *objs, "-o", str(exe), "-pthread", "-lm"]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0 and "cannot find" in res.stderr and "san" in res.stderr:
        pytest.skip("sanitizer runtime not installed")
    assert res.returncode == 0, res.stderr[-3000:]
    return exe


def run(exe: Path, *args: str) -> None:
    res = subprocess.run([str(exe), *args], capture_output=True, text=True, timeout=600)
    assert res.returncode == 0 and "ERROR: " not in res.stderr and "WARNING: ThreadSanitizer" not in res.stderr, \
        (res.stdout[-2000:], res.stderr[-4000:])
    assert "host checks: 0 failed" in res.stdout


def test_host_code_under_address_and_ub_sanitizers(tmp_path):
    exe = build(tmp_path, "address,undefined")
    run(exe, str(tmp_path))


def test_threaded_host_code_under_thread_sanitizer(tmp_path):
    exe = build(tmp_path, "thread")
    run(exe, str(tmp_path), "threads")
