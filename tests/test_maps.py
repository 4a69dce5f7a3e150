"""MovingAI maps: the loader, and the six maps the reference ships
(datasets/movingai, stored gzip-compressed under tests/golden/maps) as extra
parity inputs.

GridMap::loadMovingAI (src/grid_map.cpp:16-76) is restated twice: in Python
(htrip_amd.grids.load_movingai) and in the C++ drop-in
(htrip_amd/include/htrip/grid_map.hpp).  Both are checked against hand-worked
cases (header forms, CRLF, short and missing rows, terrain letters) and, when
oracle/_ref is built, against the reference's own loader on every shipped map.
The oracle then answers queries on a shipped map exactly as BFS does.
"""
import gzip
import os
import subprocess

import numpy as np
import pytest

from htrip_amd import grids
from oracle import oracle as orc
from oracle.oracle import Oracle, grid_bfs

HERE = os.path.dirname(os.path.abspath(__file__))
MAPDIR = os.path.join(HERE, "golden", "maps")
NAMES = sorted(f[:-len(".map.gz")] for f in os.listdir(MAPDIR) if f.endswith(".map.gz"))
F, X = 1, 2

CASES = {
    "crlf and terrain": (b"type octile\r\nheight 3\r\nwidth 4\r\nmap\r\n.G@S\r\nT..W\r\n....\r\n",
                         [[F, F, X, F], [X, F, F, X], [F, F, F, F]]),
    "short and missing rows": (b"type octile\nheight 3\nwidth 5\nmap\n..\n.....\n",
                               [[F, F, X, X, X], [F, F, F, F, F], [X, X, X, X, X]]),
    "no final newline": (b"height 2\nwidth 2\nmap\n..\n.", [[F, F], [F, X]]),
    "padded header": (b"type octile\nheight   2  \nwidth 3\t\nmap  \n...\n@O.\n", [[F, F, F], [X, X, F]]),
    "long rows": (b"height 1\nwidth 2\nmap\n.....\n", [[F, F]]),
    "width before height": (b"width 3\nheight 1\nmap\nS.G\n", [[F, F, F]]),
    "no size": (b"type octile\nmap\n...\n", None),
    "zero height": (b"height 0\nwidth 4\nmap\n....\n", None),
}


def dropin_loader(tmp_path):
    exe = str(tmp_path / "load_map")
    subprocess.run(["g++", "-std=c++20", "-O1", "-I", os.path.join(HERE, "..", "htrip_amd", "include"),
                    os.path.join(HERE, "cpp", "load_map.cpp"), "-o", exe], check=True)

    def load(path):
        res = subprocess.run([exe, path], capture_output=True)
        if res.returncode != 0:
            return None
        head, _, body = res.stdout.partition(b"\n")
        r, c = (int(v) for v in head.split())
        return np.frombuffer(body, dtype=np.uint8).reshape(r, c)

    return load


@pytest.fixture(scope="module")
def cpp_load(tmp_path_factory):
    return dropin_loader(tmp_path_factory.mktemp("dropin"))


@pytest.mark.parametrize("case", sorted(CASES))
def test_loader_cases(case, tmp_path, cpp_load):
    text, expect = CASES[case]
    path = str(tmp_path / "m.map")
    with open(path, "wb") as f:
        f.write(text)
    got_cpp = cpp_load(path)
    if expect is None:
        with pytest.raises(ValueError):
            grids.load_movingai(text)
        assert got_cpp is None
    else:
        np.testing.assert_array_equal(grids.load_movingai(text), np.array(expect, np.uint8))
        np.testing.assert_array_equal(got_cpp, np.array(expect, np.uint8))
    if orc.available("reference"):
        ref = orc.reference_load_movingai(path)
        assert (ref is None) == (expect is None)
        if ref is not None:
            np.testing.assert_array_equal(ref, np.array(expect, np.uint8))


@pytest.mark.parametrize("name", NAMES)
def test_shipped_maps_load_alike(name, tmp_path, cpp_load):
    gz = os.path.join(MAPDIR, name + ".map.gz")
    path = str(tmp_path / (name + ".map"))
    with open(path, "wb") as f:
        f.write(gzip.decompress(open(gz, "rb").read()))
    py = grids.load_movingai(gz)
    assert py.shape[0] == py.shape[1] and py.shape[0] in (256, 512, 1024)
    assert 0.2 < (py == F).mean() < 0.99
    np.testing.assert_array_equal(cpp_load(path), py)
    if orc.available("reference"):
        np.testing.assert_array_equal(orc.reference_load_movingai(path), py)


def test_oracle_on_a_shipped_map():
    """Boston_0_256 fully revealed (campaign_support.cpp:150-159), b = 16: the
    oracle's matrix equals BFS on viewpoints of the largest component."""
    truth = grids.load_movingai(os.path.join(MAPDIR, "Boston_0_256.map.gz"))
    O = Oracle("port", truth, 16, 2)
    assert O.build()
    comp = grids._largest_component(truth == F)
    vps = grids.sample_free(np.where(comp, F, X).astype(np.uint8), 40, 7)
    m = O.query(vps)
    np.testing.assert_array_equal(m, grid_bfs(truth, vps))
    assert m.max() < orc.INF
