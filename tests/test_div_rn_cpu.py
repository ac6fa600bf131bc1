"""The exact shared-divisor quotient the kernels use (div_rn, disco_hd.h):
several quotients by one divisor take one division for y = 1/b and then
q = a*y, r = fma(-b, q, a), q + r*y each.  By Markstein's theorem that is
the correctly rounded a/b; this pins it on the host build (gcc, libm's fma)
on 4 million random pairs, a third of them a rounding-boundary case, which
is what lets the leaves that use it stay bit-identical to the reference's
(tests/test_leaves_cpu.py)."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_div_rn_equals_division(tmp_path):
    exe = str(tmp_path / "div_rn_probe")
    subprocess.run(["gcc", "-O2", "-std=gnu99", "-I", os.path.join(ROOT, "disco_amd", "csrc"),
                    os.path.join(ROOT, "tests", "probes", "div_rn_probe.c"), "-o", exe, "-lm"], check=True, timeout=120)
    out = subprocess.run([exe, "4000000"], capture_output=True, text=True, timeout=300)
    print(out.stdout)
    assert out.returncode == 0, out.stdout
    assert "mismatches 0" in out.stdout
