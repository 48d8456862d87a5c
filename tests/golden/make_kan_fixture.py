"""Copy the numeric arrays of the reference's KAN-head fixture
(tests/fixtures/kan_head_init_seed42.npz: weights, 64 input rows, expected
outputs and parameter gradients of sum(n + q_spatial + p_spatial)) into
tests/golden/kan_head_seed42.npz. Arrays are read one by one without
unpickling; the fixture's one object array (not numeric) is skipped."""
import sys

import numpy as np

SRC = "/root/reference/tests/fixtures/kan_head_init_seed42.npz"
DST = "tests/golden/kan_head_seed42.npz"


def main(src=SRC, dst=DST):
    out = {}
    with np.load(src, allow_pickle=False) as d:
        for k in d.files:
            try:
                a = d[k]
            except ValueError:
                print(f"skip {k} (object array)")
                continue
            out[k] = np.ascontiguousarray(a)
    np.savez_compressed(dst, **out)
    print(f"wrote {dst}: {len(out)} arrays")


if __name__ == "__main__":
    main(*sys.argv[1:])
