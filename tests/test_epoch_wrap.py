"""The launch-epoch rule of the tagged words (ddrs_amd/csrc/ddrs_epoch.h),
compiled on the host: epoch 0 tags never-written memory, so the sequence of
launch epochs skips it when the 32-bit counter wraps."""
import ctypes
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_next_epoch_skips_zero(tmp_path):
    src = tmp_path / "epoch.cpp"
    src.write_text('#include "ddrs_epoch.h"\n'
                   'extern "C" unsigned next(unsigned e) { return ddrs::next_epoch(e); }\n')
    lib = tmp_path / "libepoch.so"
    subprocess.run(["g++", "-O2", "-shared", "-fPIC", "-I",
                    os.path.join(ROOT, "ddrs_amd", "csrc"), str(src), "-o", str(lib)],
                   check=True)
    nxt = ctypes.CDLL(str(lib)).next
    nxt.restype = ctypes.c_uint32
    nxt.argtypes = [ctypes.c_uint32]
    assert nxt(0) == 1                  # the first launch of a fresh buffer
    assert nxt(41) == 42
    assert nxt(0xFFFFFFFE) == 0xFFFFFFFF
    assert nxt(0xFFFFFFFF) == 1         # the wrap skips the never-written tag
    # a whole cycle never produces 0
    e, seen_zero = 0xFFFFFFF0, False
    for _ in range(64):
        e = nxt(e)
        seen_zero |= e == 0
    assert not seen_zero
