"""Copy the reference's Juniata network stores (examples/juniata/data/
juniata_conus_adjacency.zarr and juniata_gages_adjacency.zarr: 213 MERIT
reaches, 212 edges, real length_m / slope; zarr v3, bytes + zstd) into
tests/golden/juniata/ as data fixtures. Files are copied byte for byte; no
reference code is read or run."""
import os
import shutil

SRC = "/root/reference/examples/juniata/data"
DST = os.path.join(os.path.dirname(os.path.abspath(__file__)), "juniata")

for store in ("juniata_conus_adjacency.zarr", "juniata_gages_adjacency.zarr"):
    src, dst = os.path.join(SRC, store), os.path.join(DST, store)
    if os.path.exists(dst):
        shutil.rmtree(dst)
    shutil.copytree(src, dst, copy_function=shutil.copyfile)
    for dp, dn, fn in os.walk(dst):
        for d in dn:
            os.chmod(os.path.join(dp, d), 0o755)
        for f in fn:
            os.chmod(os.path.join(dp, f), 0o644)
    print("copied", store)
