"""Writes the logarithm table of mc_math.h (mc_ln_table.h).

ln x for a positive normal fp32 x is reduced as x = 2^k z with the bit pattern
of z in [kLnOff, kLnOff + 2^23): z lies in [0.7109, 1.4219). The 32 entries
split that range by the top five fraction bits of z - kLnOff; entry i holds
invc, a double near 1/c_i (c_i the middle of its interval), and
logc = -ln(invc) rounded to double, so that

    ln x = k ln2 + logc + ln(1 + r),  r = z invc - 1,  |r| <= 2^-6.

The entry that holds 1.0 has invc = 1 and logc = 0, so that r = z - 1 is exact
next to x = 1 and ln x keeps its relative accuracy there. logc is taken from a
60-digit decimal logarithm; Python converts it to the nearest double.

Usage: python gen_ln_table.py > mc_ln_table.h
"""
import struct
from decimal import Decimal, getcontext

NB = 5
N = 1 << NB
SHIFT = 23 - NB
# 1.0 sits in the middle (in bits) of entry 18
OFF = 0x3F800000 - (37 << (SHIFT - 1))


def f32(bits):
    return struct.unpack("<f", struct.pack("<I", bits))[0]


def entries():
    getcontext().prec = 60
    out = []
    for i in range(N):
        lo = f32(OFF + (i << SHIFT))
        hi = f32(OFF + ((i + 1) << SHIFT))
        if lo <= 1.0 < hi:
            out.append((1.0, 0.0))
            continue
        c = (Decimal(lo) + Decimal(hi)) / 2
        invc = float(1 / c)
        logc = float(-Decimal(invc).ln())
        out.append((invc, logc))
    return out


def render():
    lines = [
        "// Generated by gen_ln_table.py: the reduction table of ln_tab (mc_math.h).",
        "// {invc, logc}: invc near 1/c for the middle c of the entry's interval,",
        "// logc = -ln(invc) rounded to double; entry 18 holds 1.0 exactly.",
        "#pragma once",
        "",
        "namespace ddrs {",
        "",
        f"constexpr unsigned kLnOff = 0x{OFF:08x}u;  // bits of 0.7109375",
        f"constexpr int kLnBits = {NB};",
        f"constexpr int kLnEntries = {N};",
        "",
        "struct LnEnt {",
        "    double invc, logc;",
        "};",
        "",
        "DDRS_TABLE LnEnt kLnTab[kLnEntries] = {",
    ]
    for invc, logc in entries():
        lines.append(f"    {{{invc.hex()}, {logc.hex()}}},")
    lines += ["};", "", "}  // namespace ddrs", ""]
    return "\n".join(lines)


if __name__ == "__main__":
    print(render(), end="")
