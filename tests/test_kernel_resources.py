"""Register and scratch budgets of the hot kernels, read from the gfx950
assembly that hipcc emits for them (no GPU needed).

The book has one persistent kernel per path class; at most 256 VGPRs keeps
two waves per SIMD, and the FP32 book kernels and the FP32 snowball kernels
must not spill to scratch. A regression here halves the book's throughput
before any timing shows it."""

import pathlib
import re
import shutil
import subprocess

import pytest

ROOT = pathlib.Path(__file__).resolve().parents[1]
CSRC = ROOT / "kiyosi_amd" / "csrc"
HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"


def kernel_resources(source, tmp_path):
    asm = tmp_path / (source + ".s")
    flags = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-Xclang", "-target-feature", "-Xclang",
             "-packed-fp32-ops", "-I", str(ROOT / "include"), "--cuda-device-only", "-S"]
    subprocess.run([HIPCC, *flags, str(CSRC / source), "-o", str(asm)], check=True, capture_output=True)
    text = asm.read_text()
    out = {}
    for block in text.split("  - .agpr_count:")[1:]:
        name = re.search(r"\.name:\s+(\S+)", block).group(1)
        vgpr = int(re.search(r"\.vgpr_count:\s+(\d+)", block).group(1))
        private = int(re.search(r"\.private_segment_fixed_size:\s+(\d+)", block).group(1))
        out[name] = (vgpr, private)
    return out


@pytest.mark.skipif(not pathlib.Path(HIPCC).exists(), reason="hipcc is not installed")
def test_book_kernels_keep_two_waves_per_simd(tmp_path):
    res = kernel_resources("book.hip", tmp_path)
    # One persistent kernel per precision, tile shape (1, 2, 4 samples per
    # thread) and path class (7).
    book = {k: v for k, v in res.items() if "book_kernel" in k}
    assert len(book) == 2 * 3 * 7
    for name, (vgpr, private) in book.items():
        assert vgpr <= 256, (name, vgpr)
        if "book_kernelILb1E" in name:  # FP32: no spills, at least three waves per SIMD
            # The arithmetic Asian (class 3) calls its payoff out of line;
            # its scratch is that call's frame.
            assert private <= (64 if "ELi3EEEv" in name else 0), (name, private)
            assert vgpr <= 168, (name, vgpr)
    # The FP32 spot-Greek kernels keep four waves per SIMD; their scratch is
    # the call frame of the arithmetic Asian's strike terms, not spills.
    greeks = {k: v for k, v in res.items() if "spot_greek_kernelILb1E" in k}
    assert len(greeks) == 2
    for name, (vgpr, private) in greeks.items():
        assert vgpr <= 128 and private <= 128, (name, vgpr, private)


@pytest.mark.skipif(not pathlib.Path(HIPCC).exists(), reason="hipcc is not installed")
def test_snowball_sobol_fp32_kernels_have_no_scratch(tmp_path):
    """The snowball's FP32 Sobol kernels: the two bridge passes and the
    single-pass kernel of plans without a bridge."""
    res = kernel_resources("autocall.hip", tmp_path)
    prefixes = ("_ZN2kg15autocall_kernelILb1ELb1ELb1E", "_ZN2kg20bridge_events_kernelILb1ELb1ELb1E",
                "_ZN2kg18bridge_fill_kernelILb1ELb1ELb1E")
    for prefix in prefixes:
        name = next(k for k in res if k.startswith(prefix))
        vgpr, private = res[name]
        assert private == 0, name
        assert vgpr <= 128, (name, vgpr)  # four or more waves per SIMD
    assert all(private == 0 for name, (_, private) in res.items() if "bridge" in name)
