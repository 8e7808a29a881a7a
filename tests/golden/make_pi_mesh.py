"""Copy the reference's pi mesh (tests/data/pi: the FESOM text mesh files its
own tests hold) into tests/golden/pi_mesh as gzip files. Data only; run in a
container that has the reference checkout:

    python tests/golden/make_pi_mesh.py /root/reference/tests/data/pi
"""
import gzip
import shutil
import sys
from pathlib import Path

FILES = ["nod2d.out", "elem2d.out", "aux3d.out", "edges.out", "edge_tri.out", "edgenum.out", "nlvls.out",
         "elvls.out"]


def main(src):
    dst = Path(__file__).resolve().parent / "pi_mesh"
    dst.mkdir(exist_ok=True)
    for f in FILES:
        with open(Path(src) / f, "rb") as fi, gzip.GzipFile(dst / (f + ".gz"), "wb", mtime=0) as fo:
            shutil.copyfileobj(fi, fo)
        print(f)


if __name__ == "__main__":
    main(sys.argv[1])
