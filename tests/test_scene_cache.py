"""The checksummed scene cache (the reference's navigation / choice caches, src/nav_field.cpp:468-688,
src/exit_choice.cpp:335-445): a fresh build writes scene_<key>.bin; the next build loads it and its tables are the
fresh ones bit for bit; a file with a flipped payload byte, a cut-off file, a foreign magic or another key's file is
refused and replaced by a fresh build; another floor has another key."""
import glob
import os

import numpy as np

import swarm_hip as sh


def tables_equal(a, b):
    ta, tb = a.tables(), b.tables()
    assert ta.keys() == tb.keys()
    for k in ta:
        assert np.array_equal(ta[k].view(np.uint8), tb[k].view(np.uint8)), k
    mg_a, mg_b = a.metric_grid(), b.metric_grid()
    for x, y in zip(mg_a, mg_b):
        assert np.array_equal(np.asarray(x), np.asarray(y))


def test_cache_round_trip_and_damaged_files(tmp_path):
    d = str(tmp_path)
    fresh = sh.Scene.layout(sh.LAYOUT_ROOM)
    first = sh.Scene.layout(sh.LAYOUT_ROOM, cache_dir=d)
    assert first.info.from_cache == 0
    files = glob.glob(os.path.join(d, "scene_*.bin"))
    assert len(files) == 1 and not glob.glob(os.path.join(d, "*.tmp*"))
    f = files[0]
    loaded = sh.Scene.layout(sh.LAYOUT_ROOM, cache_dir=d)
    assert loaded.info.from_cache == 1
    tables_equal(loaded, fresh)
    tables_equal(first, fresh)
    good = open(f, "rb").read()

    def damaged(data):
        open(f, "wb").write(data)
        s = sh.Scene.layout(sh.LAYOUT_ROOM, cache_dir=d)
        assert s.info.from_cache == 0, "a damaged file must be rebuilt"
        tables_equal(s, fresh)
        assert open(f, "rb").read() == good, "the rebuilt scene rewrites the file"
        assert sh.Scene.layout(sh.LAYOUT_ROOM, cache_dir=d).info.from_cache == 1

    flipped = bytearray(good)
    flipped[len(good) // 2] ^= 0x40
    damaged(bytes(flipped))                                # a payload byte
    damaged(good[: len(good) - 1000])                      # cut off
    damaged(b"XXSCENE1" + good[8:])                        # another magic
    tail = bytearray(good)
    tail[-1] ^= 1
    damaged(bytes(tail))                                   # the checksum itself


def test_another_floor_has_another_key(tmp_path):
    d = str(tmp_path)
    sh.Scene.layout(sh.LAYOUT_ROOM, cache_dir=d)
    sh.Scene.layout(sh.LAYOUT_PILLAR, cache_dir=d)
    sh.Scene.layout(sh.LAYOUT_ROOM, cache_dir=d, max_radius=0.3 / sh.M_PER_UNIT)   # other ring bounds: another key
    assert len(glob.glob(os.path.join(d, "scene_*.bin"))) == 3
    assert sh.Scene.layout(sh.LAYOUT_PILLAR, cache_dir=d).info.from_cache == 1
