"""div_by (titan_c-_amd/csrc/fdiv.h): FP64 division through a known correctly rounded reciprocal
with Markstein's correction must give the IEEE quotient bit for bit.  The header is compiled for
the host with g++ and checked on random operands over +-300 binades, on denominators next to powers
of two, on numerators a few ulps off exact multiples, and on signed zeros."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))


def test_div_by_matches_ieee_division(tmp_path):
    exe = str(tmp_path / "fdiv_check")
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-I/opt/rocm/include", "-D__HIP_PLATFORM_AMD__",
                    os.path.join(HERE, "native", "fdiv_check.cpp"), "-o", exe], check=True)
    r = subprocess.run([exe, "4000000"], capture_output=True, text=True)
    bad, tried = map(int, r.stdout.split()[-2:])
    assert tried > 6_900_000
    assert bad == 0, r.stdout
