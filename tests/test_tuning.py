"""The launch-shape table (rama_tuning_* in include/rama_hip.h): parsing,
size buckets, defaults and errors.  Host-side only, no kernel runs; the
GPU tests run with whatever table is committed, which changes no result."""
import os

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
TABLE = os.path.join(ROOT, "rama_amd", "tuning", "gfx950.table")
KNOBS = ["sort_path_u64", "sort_path_u32", "sort_threads", "scan_chunks", "pick_blocks", "match_blocks",
         "mp_fused_max_blocks", "cycle_chunk_pct"]


@pytest.fixture
def L():
    from rama_amd import _capi

    lib = _capi.lib()
    yield lib
    # back to what the library starts with: the committed table if any
    lib.rama_tuning_load(TABLE.encode() if os.path.exists(TABLE) else None)


def test_defaults(L):
    assert L.rama_tuning_load(None) == 0
    assert L.rama_tuning_source() == b"defaults"
    # the shapes used before any table: top-digit path up to 64 tiles, then
    # three launches per pass; 1024 threads per tile; 2048 scan chunks
    assert L.rama_tuning_get(b"sort_path_u64", 64 * 4096) == 2
    assert L.rama_tuning_get(b"sort_path_u64", 64 * 4096 + 1) == 0
    assert L.rama_tuning_get(b"sort_path_u32", 10_000_000) == 0
    assert L.rama_tuning_get(b"sort_threads", 3000) == 1024
    assert L.rama_tuning_get(b"scan_chunks", 1) == 2048
    assert L.rama_tuning_get(b"pick_blocks", 1) == 2048
    assert L.rama_tuning_get(b"match_blocks", 1) == 1024
    assert L.rama_tuning_get(b"mp_fused_max_blocks", 1) == 256
    assert L.rama_tuning_get(b"cycle_chunk_pct", 1) == 125
    assert L.rama_tuning_get(b"no_such_knob", 1) == -1


def test_buckets_and_fallback(L, tmp_path):
    p = tmp_path / "t.table"
    p.write_text("# comment line\n"
                 "sort_path_u64 100000 1   # onesweep up to 100k pairs\n"
                 "sort_path_u64 -1 0\n"
                 "sort_path_u64 5000 2\n"      # out of order on purpose: rows are kept sorted
                 "\n"
                 "scan_chunks -1 1024\n")
    assert L.rama_tuning_load(str(p).encode()) == 0
    assert L.rama_tuning_source() == str(p).encode()
    assert [L.rama_tuning_get(b"sort_path_u64", n) for n in (1, 5000, 5001, 100000, 100001, 10**9)] == [2, 2, 1, 1, 0, 0]
    assert L.rama_tuning_get(b"scan_chunks", 7) == 1024
    # knobs the file leaves out keep their defaults, as do sizes past the
    # last row when it has a limit
    assert L.rama_tuning_get(b"pick_blocks", 7) == 2048
    p.write_text("match_blocks 10 4096\n")
    assert L.rama_tuning_load(str(p).encode()) == 0
    assert L.rama_tuning_get(b"match_blocks", 10) == 4096
    assert L.rama_tuning_get(b"match_blocks", 11) == 1024


@pytest.mark.parametrize("text, msg", [
    ("sort_path_u64 100 3\n", b"out of range"),
    ("sort_threads -1 768\n", b"512 or 1024"),
    ("bogus_knob -1 1\n", b"unknown knob"),
    ("scan_chunks -1\n", b"expected 'knob max_size value'"),
    ("scan_chunks -1 1024 9\n", b"expected 'knob max_size value'"),
    ("scan_chunks -7 1024\n", b"max_size"),
])
def test_rejected_tables(L, tmp_path, text, msg):
    p = tmp_path / "bad.table"
    p.write_text(text)
    L.rama_tuning_load(None)
    assert L.rama_tuning_load(str(p).encode()) == 1
    assert msg in L.rama_last_error()
    assert L.rama_tuning_source() == b"defaults"  # the table in use is unchanged


def test_set_marks_modified(L):
    assert L.rama_tuning_load(None) == 0
    assert L.rama_tuning_set(b"sort_threads", -1, 512) == 0
    assert L.rama_tuning_get(b"sort_threads", 1) == 512
    assert L.rama_tuning_source() == b"modified defaults"
    assert L.rama_tuning_set(b"nope", -1, 1) == 1


def test_committed_table_parses(L):
    if not os.path.exists(TABLE):
        pytest.skip("no tuning table committed yet")
    assert L.rama_tuning_load(TABLE.encode()) == 0, L.rama_last_error()
    names = set()
    for line in open(TABLE):
        line = line.split("#")[0].split()
        if line:
            names.add(line[0])
            assert L.rama_tuning_get(line[0].encode(), int(line[1]) if int(line[1]) >= 0 else 1 << 40) == int(line[2])
    assert names <= set(KNOBS)
