"""The rasterizer's exact fast-forward of serial float sums (rf_ff.hpp) equals
the serial loop bit for bit (host build of the same header)."""
import os
import subprocess
import tempfile

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_ff_add_matches_serial_loop():
    src = os.path.join(ROOT, "tests", "native", "ff_check.cpp")
    with tempfile.TemporaryDirectory() as d:
        exe = os.path.join(d, "ff_check")
        subprocess.check_call(["/opt/rocm/bin/hipcc", "-O2", "-std=c++17", "-ffp-contract=off", "-x", "c++",
                               "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include", src, "-o", exe])
        out = subprocess.run([exe, "300000"], capture_output=True, text=True)
    total, bad = (int(x) for x in out.stdout.split()[-2:])
    assert total > 900000
    assert bad == 0, out.stdout
