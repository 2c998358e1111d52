"""Host checks of the tiled layout (tests/cpp/test_tiles.cpp): build_tiles on
seeded random segment matrices -- every nonzero stored once at its
block-local index, whole slices per wavefront run, one partial slot per
(block, segment) pair in block order, a CPU walk of the layout reproducing
every segment's dot product, and the same layout from two builds (the
layout is cut on host threads).  Runs on the CPU; nothing of it uses a GPU."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
BIN = os.path.join(HERE, "cpp", "bin", "test_tiles")


def test_tiled_layout_on_host():
    if not os.path.exists(BIN):
        subprocess.run(["make", "-C", os.path.join(HERE, "cpp"), BIN], check=True)
    r = subprocess.run([BIN], capture_output=True, text=True, timeout=300)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "all tiled-layout checks passed" in r.stdout
