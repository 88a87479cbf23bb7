"""The angle wrap of the update kernel (qcbc_wrap.h: wrap_down / wrap_up)
against update()'s literal loops (QuickCBC.c:2336-2342), bit for bit, built
for the host from the same header with g++ -ffp-contract=off: random angles
up to 1e4 rad, log-uniform magnitudes, every binade boundary and its
neighbours, multiples of 2 pi and pi."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_wrap_equals_literal_loops(tmp_path):
    exe = tmp_path / "wrap_check"
    subprocess.run(["g++", "-O2", "-ffp-contract=off", "-std=c++17", "-I", os.path.join(ROOT, "quickcbc_amd", "csrc"),
                    "-o", str(exe), os.path.join(ROOT, "tests", "native", "wrap_check.cpp")], check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.startswith("ok ")
    assert int(r.stdout.split()[1]) > 800000
