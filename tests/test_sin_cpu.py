"""The device's slerp sine (luxcore_amd/csrc/lrh_sin.h) against the C library
the oracle uses, exhaustively: every float argument in [0, pi rounded up],
1,078,530,012 of them, must round to the same float.  The motion sample
(src/luxrays/core/geometry/quaternion.cpp:143-164 Slerp) evaluates
sin((1-t) phi) and sin(t phi) with phi in [0, pi/2], so this pins that step of
the two-level path completely."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_slerp_sine_equals_libc_for_every_float(tmp_path):
    exe = tmp_path / "sin_exhaustive"
    src = os.path.join(ROOT, "tests", "native", "sin_exhaustive.cpp")
    subprocess.run(["g++", "-O2", "-ffp-contract=off", "-pthread", "-o", str(exe), src], check=True, timeout=120)
    p = subprocess.run([str(exe), str(min(os.cpu_count() or 1, 8))], capture_output=True, text=True, timeout=600)
    print(p.stdout)
    assert p.returncode == 0, p.stdout + p.stderr
    assert "arguments 1078530012 differing 0" in p.stdout
