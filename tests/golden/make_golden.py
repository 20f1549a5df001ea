"""Writes the golden fixtures of tests/golden/ from the reference's own CPU code.

Runs oracle/_ref/ref_driver (built from /root/reference by oracle/ref_build.sh; this container only) and
packs its records into compressed .npz files:
  terrace_kat.npz  CpuCrowdSim steps on the terrace pitch with forced drive directions (state, sorted grid,
                   cell table after every step)
  nav_kat.npz      nav::buildGrid / buildExitField / buildChoiceTable on rectangle worlds
  choice_kat.npz   nav::choiceStep on the two-door room (test_gpu nav.exit_choice's known answers)
  spawn_kat.npz    seatedSpawnSeats (mt19937's raw output, partial Fisher-Yates) on venues S, M and L
  venue_kat.npz    the analytic seated venues S, M and L: buildSeatedLayout's layout, the cells of every band, the seats
                   (venue L: every 17th) and seated::sdfGrad / floorHeight / cellAt / wall / floorSlope / exitByLat /
                   gangwayAt at seeded probe points
  seated_kat.npz   CpuCrowdSim's seated 3D step on venue S (its 100 % seat spawn, 4,096 slots): a forced random drive for 40
                   steps and the default model for 200 steps (states, contacts, the choice state), with the navigation field
                   and choice table it used (checksums and every 997th cell)
  metrics_kat.npz  metrics::cicDeposit / speedMass known answers, MetricVenue::cellInfo thresholds and six steps of
                   MetricsReference::record on synthetic states (64 x 64 grid of 1 m cells)

Usage: python tests/golden/make_golden.py
"""
import os
import struct
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
DTYPES = {b"f": np.float32, b"u": np.uint32, b"i": np.int32, b"b": np.uint8, b"d": np.float64}


def read_records(path):
    out = {}
    with open(path, "rb") as f:
        data = f.read()
    pos = 0
    while pos < len(data):
        (n,) = struct.unpack_from("<I", data, pos); pos += 4
        name = data[pos:pos + n].decode(); pos += n
        dt = DTYPES[data[pos:pos + 1]]; pos += 1
        (nd,) = struct.unpack_from("<I", data, pos); pos += 4
        dims = struct.unpack_from("<%dQ" % nd, data, pos); pos += 8 * nd
        cnt = int(np.prod(dims)) if nd else 0
        arr = np.frombuffer(data, dtype=dt, count=cnt, offset=pos).reshape(dims).copy()
        pos += cnt * np.dtype(dt).itemsize
        out[name] = arr
    return out


def main():
    drv = os.path.join(REPO, "oracle", "_ref", "ref_driver")
    subprocess.run([os.path.join(REPO, "oracle", "ref_build.sh")], check=True)
    with tempfile.TemporaryDirectory() as td:
        rec = os.path.join(td, "golden.bin")
        # the reference keeps its venue navigation caches in SWARM_NAV_CACHE: a temporary folder, never the reference's tree
        env = dict(os.environ, SWARM_NAV_CACHE=os.path.join(tempfile.gettempdir(), "swarm_ref_navcache"))
        subprocess.run([drv, rec], check=True, env=env)
        r = read_records(rec)
    terrace = {k.replace("/", "__"): v for k, v in r.items() if k.startswith("t_")}
    choice = {k.replace("/", "__"): v for k, v in r.items() if k.startswith("choice_")}
    nav = {}
    for k, v in r.items():
        if not (k.startswith("room2/") or k.startswith("lcorr/") or k.startswith("room3/")):
            continue
        if k.endswith("/table/slots"):   # keep the non-empty candidate slots only (at most the exit count)
            ne = int(r[k.replace("slots", "inv_cap_s")].size)
            v = v[:, :ne]
        if k.endswith("/table/dists"):
            ne = int(r[k.replace("dists", "inv_cap_s")].size)
            v = v[:, : (ne + 1) // 2]
        nav[k.replace("/", "__")] = v
    np.savez_compressed(os.path.join(HERE, "terrace_kat.npz"), **terrace)
    np.savez_compressed(os.path.join(HERE, "nav_kat.npz"), **nav)
    np.savez_compressed(os.path.join(HERE, "choice_kat.npz"), **choice)
    metrics = {k.replace("/", "__"): v for k, v in r.items() if k.startswith("metrics/")}
    np.savez_compressed(os.path.join(HERE, "metrics_kat.npz"), **metrics)
    spawn = {k.replace("/", "__"): v for k, v in r.items() if k.startswith("spawn/")}
    np.savez_compressed(os.path.join(HERE, "spawn_kat.npz"), **spawn)
    venue = {k.replace("/", "__"): v for k, v in r.items() if k.startswith("venue/")}
    np.savez_compressed(os.path.join(HERE, "venue_kat.npz"), **venue)
    seated = {k.replace("/", "__"): v for k, v in r.items() if k.startswith("seated/")}
    np.savez_compressed(os.path.join(HERE, "seated_kat.npz"), **seated)
    for f in ("terrace_kat.npz", "nav_kat.npz", "choice_kat.npz", "metrics_kat.npz", "spawn_kat.npz", "venue_kat.npz", "seated_kat.npz"):
        print(f, os.path.getsize(os.path.join(HERE, f)), "bytes")


if __name__ == "__main__":
    sys.exit(main())
