"""The seated venues' cell table (the GPU step's cell source) answers every heightfield query as the generator does.

The GPU step reads a venue's cells from a per-band table with a bucketed lateral search, and its wall and distance
queries visit only the run of cells that can still be nearest (the table's bands are sorted). The host queries read the
generator, scanning whole bands as the reference does. tests/venue_table_check.cpp evaluates both on the CPU for venues
S, M and L at uniform points at the venues' heights, on and next to every cell and band boundary (the nearest-cell
searches' ties; these also through a table whose lateral buckets all start at the band's last cell, so that every
search walks back), on and next to the table's bucket edges, and at non-finite points: cell key, floor height, kind,
signed distance and gradient, the nearest wall at four reaches, the incline slope, the gangway and the step-up scan must
agree bit for bit; and the band lookup (its row estimated by a product with the reciprocal row depth) must give the
band that the reference's division gives, at 64 floats either side of every row front and clearway edge and at a
million random depths per venue. The program is compiled on the host and linked against libswarm_hip.so (nothing runs on a GPU).
"""
import os
import shutil
import subprocess

import pytest

import swarm_hip as sh

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HIPCC = "/opt/rocm/bin/hipcc"


@pytest.mark.skipif(not os.path.exists(HIPCC) and shutil.which("hipcc") is None, reason="hipcc is absent")
def test_cell_table_queries_match_generator(tmp_path):
    exe = tmp_path / "venue_table_check"
    lib_dir = os.path.dirname(sh.LIB_PATH)
    hipcc = HIPCC if os.path.exists(HIPCC) else shutil.which("hipcc")
    cmd = [hipcc, "-std=c++17", "-O2", "-ffp-contract=off", "-I", os.path.join(lib_dir, "csrc"), "-I", os.path.join(REPO, "include"),
           os.path.join(REPO, "tests", "venue_table_check.cpp"), "-o", str(exe), "-L", lib_dir, "-lswarm_hip", "-Wl,-rpath," + lib_dir]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-4000:]
    r = subprocess.run([str(exe), "100000"], capture_output=True, text=True, timeout=600)
    assert r.returncode == 0 and r.stdout.startswith("ok"), r.stdout + r.stderr
