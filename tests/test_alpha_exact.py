"""The closed form of the Siddon alpha accumulators (apirl_amd/csrc/alpha_exact.h),
used by the gather backprojector, equals the reference's repeated float additions
(recon/src/Siddon.cpp:322,330,338) bit for bit: brute force over random starts,
increments and step counts, a fifth of them forced onto round-half-even ties."""
import os
import subprocess
import tempfile

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_alpha_after_equals_repeated_addition():
    with tempfile.TemporaryDirectory() as td:
        exe = os.path.join(td, "aec")
        subprocess.run(["g++", "-O2", "-ffp-contract=off", "-std=c++17",
                        "-I", os.path.join(ROOT, "apirl_amd", "csrc"),
                        os.path.join(ROOT, "tests", "native", "alpha_exact_check.cpp"), "-o", exe],
                       check=True)
        out = subprocess.run([exe, "1000000"], capture_output=True, text=True)
        assert out.returncode == 0, out.stdout
        assert "mismatches=0" in out.stdout
