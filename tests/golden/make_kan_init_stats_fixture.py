"""Copy the reference's DDR init statistics (tests/fixtures/kan_init_stats_ddr.csv:
per-tensor shape, mean, std, min, max and mean |x| of a freshly initialised
DDR KAN head with 10 attributes, 3 outputs, hidden 21, 2 blocks, grid 50,
order 2) into tests/golden/kan_init_stats_ddr.csv. The rows are data; only
the columns are checked."""
import csv
import sys

SRC = "/root/reference/tests/fixtures/kan_init_stats_ddr.csv"
DST = "tests/golden/kan_init_stats_ddr.csv"
COLS = ("name", "shape", "mean", "std", "min", "max", "abs_mean")


def main(src=SRC, dst=DST):
    with open(src, newline="") as f:
        rows = list(csv.DictReader(f))
    assert rows and tuple(rows[0].keys()) == COLS, rows[0].keys() if rows else None
    with open(dst, "w", newline="") as f:
        w = csv.DictWriter(f, fieldnames=COLS)
        w.writeheader()
        for r in rows:
            w.writerow({c: r[c] for c in COLS})
    print(f"wrote {dst}: {len(rows)} rows")


if __name__ == "__main__":
    main(*sys.argv[1:])
