"""The launch-epoch rule of the tagged words (ddrs_amd/csrc/ddrs_epoch.h),
compiled on the host: tag 0 marks never-written memory, so the sequence of
launch epochs skips every epoch whose word or row tag could be 0 (0 and 2^31),
and the row tags of adjacent rows differ."""
import ctypes
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_next_epoch_skips_zero(tmp_path):
    src = tmp_path / "epoch.cpp"
    src.write_text('#include "ddrs_epoch.h"\n'
                   'extern "C" unsigned next(unsigned e) { return ddrs::next_epoch(e); }\n'
                   'extern "C" unsigned rtag(unsigned e, int r) { return ddrs::row_tag(e, r); }\n')
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
    assert nxt(0x7FFFFFFF) == 0x80000001  # 2^31: its odd rows would tag 0
    rtag = ctypes.CDLL(str(lib)).rtag
    rtag.restype = ctypes.c_uint32
    rtag.argtypes = [ctypes.c_uint32, ctypes.c_int]
    assert rtag(5, 0) == 5 and rtag(5, 1) == 0x80000005 and rtag(5, 2) == 5
    # across both skips, no epoch and no row tag is 0, and adjacent rows differ
    for start in (0xFFFFFFF0, 0x7FFFFFF0):
        e = start
        for _ in range(64):
            e = nxt(e)
            assert e != 0
            for r in range(4):
                assert rtag(e, r) != 0 and rtag(e, r) != rtag(e, r + 1)
