"""The launch-shape tuning table (vgpu_tune_*): global knobs, cardinality
buckets and the autotuner's file format.  Host-side only."""
import pytest


@pytest.fixture
def lib(native):
    lib = native.vgpu()
    yield lib
    lib.vgpu_tune_reset()
    table = native.PKG / "tuning.txt"
    if table.exists():
        assert lib.vgpu_tune_load(str(table).encode()) >= 0


def test_buckets_pick_the_smallest_covering_one(lib):
    lib.vgpu_tune_reset()
    assert lib.vgpu_tune_get(b"gb.lds.max_blocks") == 1024
    assert lib.vgpu_tune_set_for(64, b"gb.lds.max_blocks", 512) == 0
    assert lib.vgpu_tune_set_for(4096, b"gb.lds.max_blocks", 2048) == 0
    assert lib.vgpu_tune_set_for(4096, b"gb.lds.rows_per_block", 4096) == 0
    assert lib.vgpu_tune_get_for(10, b"gb.lds.max_blocks") == 512
    assert lib.vgpu_tune_get_for(64, b"gb.lds.max_blocks") == 512
    assert lib.vgpu_tune_get_for(65, b"gb.lds.max_blocks") == 2048
    assert lib.vgpu_tune_get_for(100000, b"gb.lds.max_blocks") == 1024          # no bucket: global
    assert lib.vgpu_tune_get_for(10, b"gb.lds.rows_per_block") == 4096           # bucket 64 leaves it: the next sets it
    assert lib.vgpu_tune_get_for(1000, b"gb.lds.rows_per_block") == 4096
    assert lib.vgpu_tune_get_for(0, b"gb.lds.max_blocks") == 1024               # unknown group count: global
    lib.vgpu_tune_reset()
    assert lib.vgpu_tune_get_for(10, b"gb.lds.max_blocks") == 1024


def test_buckets_by_width_and_rows(lib):
    """a bucket covers groups, record width and rows; the first covering one
    that sets the knob wins, narrowest width class first, then the smallest
    row class, then the smallest cardinality"""
    lib.vgpu_tune_reset()
    big = 1 << 62
    assert lib.vgpu_tune_set_at(1024, 4, big, b"gb.lds.max_blocks", 256) == 0          # narrow records
    assert lib.vgpu_tune_set_at(1024, big, 4_000_000, b"gb.lds.max_blocks", 512) == 0  # short launches
    assert lib.vgpu_tune_set_at(1024, big, big, b"gb.lds.max_blocks", 2048) == 0       # the rest
    assert lib.vgpu_tune_get_at(100, 3, 50_000_000, b"gb.lds.max_blocks") == 256
    assert lib.vgpu_tune_get_at(100, 14, 1_000_000, b"gb.lds.max_blocks") == 512
    assert lib.vgpu_tune_get_at(100, 14, 50_000_000, b"gb.lds.max_blocks") == 2048
    assert lib.vgpu_tune_get_at(5000, 3, 1_000, b"gb.lds.max_blocks") == 1024          # beyond 1024 groups: global
    assert lib.vgpu_tune_get_at(100, 0, 0, b"gb.lds.max_blocks") == 256                # unknown width / rows
    assert lib.vgpu_tune_get_for(100, b"gb.lds.max_blocks") == 256
    assert lib.vgpu_tune_set_at(1024, 0, big, b"gb.lds.max_blocks", 8) == -1
    assert lib.vgpu_tune_set_at(1 << 20, big, big, b"gb.glb.load_pct", 70) == 0
    assert lib.vgpu_tune_get_at(100_000, 9, 10, b"gb.glb.load_pct") == 70
    assert lib.vgpu_tune_set(b"gb.glb.load_pct", 95) == -1 and lib.vgpu_tune_set(b"gb.glb.load_pct", 20) == -1
    assert lib.vgpu_tune_set(b"join.load_pct", 96) == -1


def test_rejects_bad_settings(lib):
    assert lib.vgpu_tune_set(b"no.such.knob", 1) == -1
    assert lib.vgpu_tune_set(b"gb.lds.max_blocks", 0) == -1
    assert lib.vgpu_tune_set(b"gb.lds.budget", 70000) == -1
    assert lib.vgpu_tune_set(b"gb.lds.slots_per_group", 1) == -1
    assert lib.vgpu_tune_set_for(0, b"gb.lds.max_blocks", 8) == -1
    assert lib.vgpu_tune_get_for(10, b"no.such.knob") == -1


def test_load_file(lib, tmp_path):
    f = tmp_path / "t.txt"
    f.write_text("# comment\n0 gb.star.items 8\n256 gb.lds.slots_per_group 4   # trailing comment\n\n"
                 "1048576 gb.part.min_groups 16384\n")
    lib.vgpu_tune_reset()
    assert lib.vgpu_tune_load(str(f).encode()) == 3
    assert lib.vgpu_tune_get(b"gb.star.items") == 8
    assert lib.vgpu_tune_get_for(100, b"gb.lds.slots_per_group") == 4
    assert lib.vgpu_tune_get_for(1000, b"gb.lds.slots_per_group") == 2
    assert lib.vgpu_tune_get_for(500000, b"gb.part.min_groups") == 16384
    bad = tmp_path / "bad.txt"
    bad.write_text("12 gb.star.items\n")
    assert lib.vgpu_tune_load(str(bad).encode()) == -1
    axes = tmp_path / "axes.txt"
    axes.write_text("1024 gb.lds.copies 4 w<=6 rows<=8000000   # narrow, short\n1024 gb.lds.copies 1\n"
                    "2000000 join.load_pct 70 rows<=2000000\n")
    lib.vgpu_tune_reset()
    assert lib.vgpu_tune_load(str(axes).encode()) == 3
    assert lib.vgpu_tune_get_at(16, 3, 1_000_000, b"gb.lds.copies") == 4
    assert lib.vgpu_tune_get_at(16, 14, 1_000_000, b"gb.lds.copies") == 1
    assert lib.vgpu_tune_get_at(16, 3, 50_000_000, b"gb.lds.copies") == 1
    assert lib.vgpu_tune_get_at(1_000_000, 4, 1_000_000, b"join.load_pct") == 70
    assert lib.vgpu_tune_get_at(10_000_000, 4, 10_000_000, b"join.load_pct") == 50
    unknown = tmp_path / "unknown.txt"
    unknown.write_text("64 gb.lds.copies 2 x<=3\n")
    assert lib.vgpu_tune_load(str(unknown).encode()) == -1
    assert lib.vgpu_tune_load(b"/nonexistent/tuning.txt") == -1
