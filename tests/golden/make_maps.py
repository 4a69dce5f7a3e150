"""Writes tests/golden/maps/*.map.gz: the six MovingAI maps the reference ships
(datasets/movingai/{street,maze}/maps), gzip-compressed, as parity inputs.
They are data (benchmark maps), stored byte for byte; the loader under test
parses them.  Run once in a checkout that has the reference:

    python tests/golden/make_maps.py [/root/reference]
"""
import gzip
import os
import sys

MAPS = ["street/maps/Boston_0_256.map", "street/maps/Paris_0_512.map", "street/maps/Boston_0_1024.map",
        "street/maps/Berlin_0_1024.map", "maze/maps/maze512-1-0.map", "maze/maps/maze512-16-0.map"]


def main(ref: str) -> None:
    here = os.path.join(os.path.dirname(os.path.abspath(__file__)), "maps")
    os.makedirs(here, exist_ok=True)
    for rel in MAPS:
        with open(os.path.join(ref, "datasets", "movingai", rel), "rb") as f:
            data = f.read()
        out = os.path.join(here, os.path.basename(rel) + ".gz")
        with open(out, "wb") as f:
            f.write(gzip.compress(data, compresslevel=9, mtime=0))
        print(out, len(data), "->", os.path.getsize(out))


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
