"""The C host (disco_amd/host/disco_main.c: the reference's main loop over
the C ABI) compiles as C99 against include/ alone and links against the
library; the header is self-contained and every symbol it declares is
exported (tests/test_oracle_cpu.py checks the exports).  tests/
test_gpu_host.py runs the linked host on the GPU."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def build_host(out: str) -> None:
    lib = os.path.join(ROOT, "disco_amd", "_lib")
    subprocess.run(["gcc", "-std=c99", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "disco_amd", "host", "disco_main.c"), "-L", lib, "-ldisco_gpu",
                    f"-Wl,-rpath,{lib}", "-o", out], check=True, timeout=120)


def test_header_alone_compiles_as_c(tmp_path):
    src = tmp_path / "h.c"
    src.write_text('#include "disco_gpu.h"\nint main(void) { disco_params p; (void)p; return 0; }\n')
    subprocess.run(["gcc", "-std=c99", "-Wall", "-Werror", "-pedantic", "-I", os.path.join(ROOT, "include"),
                    "-c", str(src), "-o", str(tmp_path / "h.o")], check=True, timeout=60)


def test_host_compiles_and_links(tmp_path):
    exe = str(tmp_path / "disco_main")
    build_host(exe)
    assert os.path.getsize(exe) > 0
