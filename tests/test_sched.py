"""Admission policy of the request loop, compiled and run on the host."""
import os
import subprocess

from conftest import ROOT


def test_admission_policy(tmp_path):
    src = os.path.join(ROOT, "tests", "native", "sched_test.cpp")
    inc = os.path.join(ROOT, "aapm_lowdosectimagerecon_competition_amd", "csrc")
    exe = tmp_path / "sched_test"
    subprocess.run(["g++", "-std=c++17", "-O1", "-fsanitize=address,undefined", "-I", inc, src,
                    "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout
    assert "sched ok" in out
