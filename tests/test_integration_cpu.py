"""The reference-side binding shown in INTEGRATION.md section 2 compiles
against LuxRays' interfaces for this path (restated with their exact virtual
signatures in luxcore_amd/host/luxrays_bases.h): every overriding member is
marked override and matches, and the device and kernel classes are concrete.
A planted mistake -- the non-const buffer argument of EnqueueReadBuffer, which
hides instead of overriding hardwaredevice.h:145 -- must not compile."""
import os
import re
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def binding_source():
    with open(os.path.join(ROOT, "INTEGRATION.md")) as f:
        text = f.read()
    sec = text[text.index("## 2. Reference-side binding"):]
    code = re.search(r"```cpp\n(.*?)```", sec, re.S).group(1)
    # the reference's headers are stood for by the restated interfaces
    code = "\n".join(ln for ln in code.splitlines()
                     if not re.match(r'\s*#include "(luxrays/|lrhip\.h)', ln))
    head = (f'#include <type_traits>\n#include "{ROOT}/luxcore_amd/host/luxrays_bases.h"\n'
            f'#include "{ROOT}/include/lrhip.h"\n')
    tail = ("\nstatic_assert(!std::is_abstract<luxrays::HIPIntersectionDevice>::value, \"device is concrete\");\n"
            "static_assert(!std::is_abstract<luxrays::HIPBVHKernel>::value, \"kernel is concrete\");\n"
            "static_assert(!std::is_abstract<luxrays::HIPDeviceBuffer>::value, \"buffer is concrete\");\n")
    return head + code + tail


def _compile(tmp_path, src):
    f = tmp_path / "binding.cpp"
    f.write_text(src)
    return subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-Dluxrays=luxrays_hip", "-Wall",
                           "-Werror=suggest-override", "-Werror=overloaded-virtual", str(f)],
                          capture_output=True, text=True, timeout=120)


def test_integration_binding_compiles_against_the_interfaces(tmp_path):
    p = _compile(tmp_path, binding_source())
    assert p.returncode == 0, p.stderr


def test_a_signature_mismatch_does_not_compile(tmp_path):
    src = binding_source()
    old = "void EnqueueReadBuffer(const HardwareDeviceBuffer *buff"
    assert old in src
    p = _compile(tmp_path, src.replace(old, "void EnqueueReadBuffer(HardwareDeviceBuffer *buff"))
    assert p.returncode != 0 and "override" in p.stderr
