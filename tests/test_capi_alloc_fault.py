"""No C++ exception crosses the C ABI when host allocations fail.

tests/native/alloc_fault.cpp replaces operator new with one that throws
std::bad_alloc while armed and calls entry points whose error paths
(astc_encode, astc_decode, astc_output_size with a rejected description) or
bodies (astc_request_bytes, astc_write_file's error message) allocate.  Each
call must return its error code with the message recorded, never terminate the
process.  Host only: no call reaches a kernel launch.
"""
import os
import subprocess

from astc_encoder_amd import _capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_error_paths_survive_failing_host_allocations(tmp_path):
    exe = tmp_path / "alloc_fault"
    libdir = os.path.dirname(_capi.LIB_PATH)
    subprocess.run(["g++", "-std=c++17", "-O1", "-I", os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "native", "alloc_fault.cpp"), "-o", str(exe), "-L", libdir,
                    "-lastc_mi355x", f"-Wl,-rpath,{libdir}", "-Wl,-rpath-link,/opt/rocm/lib"], check=True)
    run = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60)
    assert run.returncode == 0, run.stderr
    out = {}
    for ln in run.stdout.splitlines():
        name, rc, msg = ln.split(" ", 2)
        out[name] = (int(rc), msg)
    # rejected arguments: the code and the full message, recorded without allocating
    for name in ("encode_bad_desc", "decode_bad_size", "output_size_bad_desc"):
        assert out[name] == (-1, "width/height must be in 1..16777215"), (name, out[name])
    # a body that allocates: the failure is caught and reported as ASTC_ERR_NO_MEMORY
    assert out["request_bytes"] == (-3, "host allocation failed")
    assert out["write_file_bad_path"][0] == -3
    # and the library works normally once allocations succeed again
    assert out["request_bytes_unarmed"][0] > 0
