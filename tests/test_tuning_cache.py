"""The autotuner's per-device, per-scene-class cache as text
(rd_tuning_cache_save / rd_tuning_cache_load): host code, no GPU."""
import pytest

from radiantium_amd import native
from radiantium_amd.native import RadiantiumError


LINES = [
    "gfx950/256|path/2|inst0|prims2^20|paths2^26\t3\t2\t2\t1\t15\t65536\t0\t1\t2\t262144\t0\t0\t8\t0\t0",
    "gfx950/256|vol_path/0|inst0|prims2^4|paths2^18\t0\t2\t2\t0\t15\t65536\t0\t0\t2\t65536\t0\t24\t16\t40\t0",
]


def test_cache_roundtrip(tmp_path):
    native.tuning_cache_clear()
    src = tmp_path / "in.txt"
    src.write_text("\n".join(LINES) + "\n")
    native.tuning_cache_load(src)
    out = tmp_path / "out.txt"
    native.tuning_cache_save(out)
    assert sorted(out.read_text().strip().splitlines()) == sorted(LINES)
    # load merges; a second load of the same keys replaces them
    native.tuning_cache_load(src)
    native.tuning_cache_save(out)
    assert len(out.read_text().strip().splitlines()) == 2
    native.tuning_cache_clear()
    native.tuning_cache_save(out)
    assert out.read_text() == ""


def test_cache_rejects_bad_files(tmp_path):
    native.tuning_cache_clear()
    bad = tmp_path / "bad.txt"
    bad.write_text("no tab here\n")
    with pytest.raises(RadiantiumError):
        native.tuning_cache_load(bad)
    bad.write_text("key\t1\t2\n")   # too few fields
    with pytest.raises(RadiantiumError):
        native.tuning_cache_load(bad)
    with pytest.raises(RadiantiumError):
        native.tuning_cache_load(tmp_path / "missing.txt")
    native.tuning_cache_save(tmp_path / "empty.txt")
    assert (tmp_path / "empty.txt").read_text() == ""   # a failed load leaves the cache as it was
