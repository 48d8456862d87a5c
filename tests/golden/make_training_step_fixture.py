"""Copy the numeric arrays of the reference's training-step fixtures
(tests/fixtures/training_step/*_77006074.npz, dumped from DDR for gauge
10336740, 19 reaches, a 2136-hour window, tau 3, warmup 5 days) into
tests/golden/training_step_77006074.npz, each array prefixed with its file's
stem (mc_forward_, daily_q_, loss_and_grads_, adam_step_). The forward's
2136 x 19 forcing is left out (its routing needs reach lengths and slopes from
a CONUS store that the fixtures do not hold). Arrays are read without
unpickling."""
import os
import sys

import numpy as np

SRC = "/root/reference/tests/fixtures/training_step"
DST = "tests/golden/training_step_77006074.npz"
FILES = ("mc_forward", "daily_q", "loss_and_grads", "adam_step")
SKIP = {"mc_forward_q_prime_full"}


def main(src=SRC, dst=DST):
    out = {}
    for stem in FILES:
        with np.load(os.path.join(src, f"{stem}_77006074.npz"), allow_pickle=False) as d:
            for k in d.files:
                name = f"{stem}_{k}"
                if name not in SKIP:
                    out[name] = np.ascontiguousarray(d[k])
    np.savez_compressed(dst, **out)
    print(f"wrote {dst}: {len(out)} arrays")


if __name__ == "__main__":
    main(*sys.argv[1:])
