"""The FFT pass kernel built for the CPU (no GPU needed).

The source of k_fft_pass (eisenjoch_amd/csrc/ep_fft.hip) is compiled with
the host compiler and AddressSanitizer and run with one std::thread per GPU
thread and a std::barrier for each __syncthreads, on buffers of exactly the
sizes the kernel may touch: every supported row length, row counts below
and above one block, every transform kind, the Poisson epilogue and the
interleaved store.  Any access outside a buffer aborts the run; the outputs
must equal direct O(n^2) transforms to 2e-6 of their largest value.
"""
import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
SRC = ROOT / "eisenjoch_amd" / "csrc" / "ep_fft.hip"
MAIN = Path(__file__).resolve().parent / "emu" / "fft_emu_main.cpp"
CXX = shutil.which("clang++") or "/opt/rocm/llvm/bin/clang++"


@pytest.fixture(scope="module")
def emu(tmp_path_factory):
    d = tmp_path_factory.mktemp("fft_emu")
    s = SRC.read_text()
    body = s[s.index("typedef float f32x2"):s.index("int fft_rows_per_block(int n)")]
    (d / "fft_kernel_extract.inc").write_text(body)
    exe = d / "fft_emu"
    r = subprocess.run([CXX, "-std=c++20", "-O1", "-g", "-fsanitize=address", "-fno-omit-frame-pointer",
                        f"-I{d}", str(MAIN), "-o", str(exe), "-lpthread"],
                       capture_output=True, text=True)
    if r.returncode != 0:
        pytest.fail(f"CPU build of the FFT kernel failed:\n{r.stderr[-3000:]}")
    return exe


@pytest.mark.parametrize("n", [16, 32, 64, 128, 256, 512])
def test_fft_kernel_on_cpu(emu, n):
    r = subprocess.run([str(emu), str(n)], capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-3000:]
    assert "emulation ok" in r.stdout
