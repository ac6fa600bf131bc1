"""The zero-gradient ghost fill the GPU kernels compile (one thread per ghost
cell, step_cells.h:bc_copy_cell) against set_cells_copy_distant's sequential
loop (boundary_functions.c:90-159) as the host runs it (host_setup.cpp:
copy_distant, itself pinned bit for bit to the reference's step-0 ghosts by
tests/test_oracle_cpu.py::test_host_setup_is_bit_exact), on random meshes,
on the identical fronts of two z-neighbours and on fronts a few ulps apart,
where the cell that starts each sum matters.  Both sides are built with the
host compiler without fused multiply-adds: every ghost value must agree to
the bit."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "disco_amd", "csrc")


def test_per_cell_fill_equals_sequential_loop(tmp_path):
    exe = str(tmp_path / "bc_probe")
    subprocess.run(["g++", "-O2", "-std=c++17", "-w", "-I", CSRC, "-I", os.path.join(ROOT, "include"), os.path.join(ROOT, "tests", "probes", "bc_probe.cpp"),
                    os.path.join(CSRC, "host_setup.cpp"), "-o", exe], check=True, timeout=300)
    out = subprocess.run([exe, "11", "600"], capture_output=True, text=True, timeout=120)
    rows = [list(map(int, line.split())) for line in out.stdout.strip().splitlines()]
    assert len(rows) == 600
    assert {r[0] for r in rows} == {0, 1, 2}
    bad = [r for r in rows if r[3]]
    assert out.returncode == 0 and not bad, bad[:5]
