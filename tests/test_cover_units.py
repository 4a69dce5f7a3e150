"""The update path's cover-product work planner (htg::plan_cover_units, host
code in libhtrip_gpu.so; no GPU call): random levels of nodes, both hops,
checked by tests/cpp/cover_units.cpp for an exact tiling of every node's tiles
and B list, the unit budget and the row-tile choice."""
import os
import subprocess

import pytest

from htrip_amd import _lib

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)


@pytest.fixture(scope="module")
def planner(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("cover_units") / "cover_units")
    libdir = os.path.dirname(_lib.LIB_PATH)
    subprocess.run(["g++", "-std=c++20", "-O1", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include",
                    "-I", os.path.join(REPO, "htrip_amd", "csrc"), os.path.join(HERE, "cpp", "cover_units.cpp"),
                    "-o", exe, "-L", libdir, "-lhtrip_gpu", f"-Wl,-rpath,{libdir}"], check=True)
    return exe


@pytest.mark.parametrize("seed", [1, 2, 3, 4])
def test_cover_units_tile_every_node_exactly(planner, seed):
    out = subprocess.run([planner, str(seed), "300"], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0 and out.stdout.startswith("ok"), out.stdout + out.stderr
