"""The host BVH builder (binned SAH with spatial splits, vulkanexp_amd/csrc/bvh.cpp) on the CPU:
tests/bvh_check.cpp is compiled with g++ against bvh.cpp and scene_io.cpp and run on seeded scenes.
It checks the tree's structure (every node reached once, leaf ranges inside the slots, depth within
the traversal's limit), that every slot holds its own triangle and a piece box inside the
triangle's box, and that random rays find through the tree the same nearest hit and any-hit
decision as a loop over every triangle (the same double-precision test on both sides)."""
import os
import subprocess

import pytest

from vulkanexp_amd import scenes

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "vulkanexp_amd", "csrc")


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("bvh") / "bvh_check")
    cmd = ["g++", "-O2", "-std=c++17", "-o", out, os.path.join(ROOT, "tests", "bvh_check.cpp"),
           os.path.join(CSRC, "bvh.cpp"), os.path.join(CSRC, "scene_io.cpp"), "-I" + os.path.join(ROOT, "include")]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    return out


SCENES = {
    "room": lambda: scenes.config_small(seed=1)[0],
    "room_open": lambda: scenes.config_small(seed=4, open_top=True)[0],
    "clustered_24": lambda: scenes.clustered_scene(seed=5, n_instances=24),
}


@pytest.mark.parametrize("name", sorted(SCENES))
def test_bvh_structure_pieces_and_hits(checker, tmp_path, name):
    path = str(tmp_path / f"{name}.scene")
    scenes.save(SCENES[name](), path)
    r = subprocess.run([checker, path, "3000", "17"], capture_output=True, text=True, timeout=600)
    print(r.stdout, r.stderr)
    assert r.returncode == 0, r.stdout + r.stderr
