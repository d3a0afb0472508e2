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
    assert lib.vgpu_tune_get_for(10, b"gb.lds.rows_per_block") == 8192           # bucket 64 leaves it global
    assert lib.vgpu_tune_get_for(1000, b"gb.lds.rows_per_block") == 4096
    assert lib.vgpu_tune_get_for(0, b"gb.lds.max_blocks") == 1024               # unknown group count: global
    lib.vgpu_tune_reset()
    assert lib.vgpu_tune_get_for(10, b"gb.lds.max_blocks") == 1024


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
    assert lib.vgpu_tune_load(b"/nonexistent/tuning.txt") == -1
