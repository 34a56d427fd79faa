"""The device's slerp sine (luxcore_amd/csrc/lrh_sinf.h) against the host C
library's sinf, which the reference's Slerp calls
(src/luxrays/core/geometry/quaternion.cpp:151-155), exhaustively: every
float in [0, 120) and in [-pi/2, 0), 2,193,166,300 arguments, must give the
same float bit for bit.  The motion sample evaluates sinf((1-t) phi) and
sinf(t phi) with phi in [0, pi/2], so this pins that step of the two-level
path completely.  Planted mistakes: a plain build of the same routine (no
fused multiply-adds) differs on arguments beyond pi/2, and a sine polynomial
with a wrong coefficient sign differs almost everywhere."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "native", "sinf_exhaustive.cpp")
HDR = os.path.join(ROOT, "luxcore_amd", "csrc", "lrh_sinf.h")


def _run(tmp_path, header=None):
    exe = tmp_path / ("sinf_exhaustive" if header is None else "sinf_planted")
    src = SRC
    if header is not None:
        src = tmp_path / "sinf_planted.cpp"
        with open(SRC) as f:
            src.write_text(f.read().replace("../../luxcore_amd/csrc/lrh_sinf.h", str(header)))
    subprocess.run(["g++", "-O2", "-ffp-contract=off", "-pthread", "-o", str(exe), str(src)], check=True,
                   timeout=120)
    p = subprocess.run([str(exe), str(min(os.cpu_count() or 1, 8))], capture_output=True, text=True, timeout=600)
    print(p.stdout)
    return p


def test_slerp_sine_equals_libc_sinf_for_every_float(tmp_path):
    p = _run(tmp_path)
    assert p.returncode == 0, p.stdout + p.stderr
    assert "arguments 2193166300 differing 0 (within pi/2: 0)" in p.stdout


def _planted(tmp_path, old, new):
    with open(HDR) as f:
        text = f.read()
    assert old in text
    h = tmp_path / "lrh_sinf_planted.h"
    h.write_text(text.replace(old, new))
    return _run(tmp_path, h)


def test_planted_mistakes_are_caught(tmp_path):
    # no fused multiply-adds (the library's build without FMA): beyond pi/2 only
    p = _planted(tmp_path, "return std::fma(a, b, c);", "return a * b + c;")
    assert p.returncode == 1 and "(within pi/2: 0)" in p.stdout
    # a wrong sign of the cubic coefficient
    p = _planted(tmp_path, "constexpr double kS1 = -0x1.555545995a603p-3;", "constexpr double kS1 = 0x1.555545995a603p-3;")
    assert p.returncode == 1 and "(within pi/2: 0)" not in p.stdout
