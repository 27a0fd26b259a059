"""Copies the reference's 2D generator output (the only data fixture the reference
holds: demos/SPGrid_Create_Simple_Domain/Create_2D_Domain/output/{size,block_offset,
flags,data}) into tests/golden/spgrid2d/ as gzip files. Run where /root/reference exists."""
import gzip
import os
import shutil

SRC = "/root/reference/demos/SPGrid_Create_Simple_Domain/Create_2D_Domain/output"
DST = os.path.join(os.path.dirname(os.path.abspath(__file__)), "spgrid2d")

if __name__ == "__main__":
    os.makedirs(DST, exist_ok=True)
    for name in ("size", "block_offset", "flags", "data"):
        with open(os.path.join(SRC, name), "rb") as f, gzip.GzipFile(os.path.join(DST, name + ".gz"), "wb", mtime=0) as g:
            shutil.copyfileobj(f, g)
    print("wrote", sorted(os.listdir(DST)))
