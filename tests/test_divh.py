"""The kernels divide by grid spacings and Re as a reciprocal multiply with
one FMA correction (divh in pllc_dev.h).  This checks on the host, with the
same operation sequence, that the result is the IEEE quotient bit for bit
(Markstein's theorem for a correctly rounded reciprocal) over random
numerators of 81 binades, the spacings of the test and benchmark grids,
their halves (the ENO second differences) and random divisors."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))


def test_reciprocal_fma_division_is_correctly_rounded(tmp_path):
    exe = tmp_path / "divh_check"
    subprocess.run(["gcc", "-O2", "-ffp-contract=off", "-o", str(exe),
                    os.path.join(HERE, "native", "divh_check.c"), "-lm"], check=True)
    out = subprocess.run([str(exe), "200000"], capture_output=True, text=True, check=True).stdout
    assert int(out.strip()) == 0
