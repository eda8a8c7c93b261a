"""The tuning table of mlcp_amd.autotune (host logic; no GPU)."""
import pytest

from mlcp_amd import autotune


def test_table_round_trip(tmp_path):
    p = str(tmp_path / "t.txt")
    assert autotune.read_table(p) == {}
    autotune.write_entry(p, "gfx950/D3/n5.6.4/B4/R3.6.18/o9/t7", (128, 2, 8))
    autotune.write_entry(p, "gfx950/D2/n4.4/B2/R2.2.4/o3/t3", (64, 1, 1))
    autotune.write_entry(p, "gfx950/D3/n5.6.4/B4/R3.6.18/o9/t7", (256, 4, 1))
    t = autotune.read_table(p)
    assert t == {"gfx950/D3/n5.6.4/B4/R3.6.18/o9/t7": (256, 4, 1),
                 "gfx950/D2/n4.4/B2/R2.2.4/o3/t3": (64, 1, 1)}
    lines = [l for l in open(p) if not l.startswith("#")]
    assert lines == sorted(lines)   # one line per key, sorted


def test_table_rejects_bad_lines(tmp_path):
    p = tmp_path / "t.txt"
    p.write_text("# c\nkey 64 1\n")
    with pytest.raises(ValueError):
        autotune.read_table(str(p))
    with pytest.raises(ValueError):
        autotune.write_entry(str(tmp_path / "u.txt"), "a b", (64, 1, 1))


def test_candidates_respect_the_lds_bound():
    assert autotune.candidates(256) == [
        (64, 1), (64, 2), (64, 4), (64, 8), (64, 16), (128, 1), (128, 2),
        (128, 4), (128, 8), (192, 1), (192, 2), (192, 4), (256, 1), (256, 2),
        (256, 4)]
    assert autotune.candidates(128) == [
        (64, 1), (64, 2), (64, 4), (64, 8), (64, 16), (128, 1), (128, 2),
        (128, 4), (128, 8)]
    assert autotune.candidates(0) == []


def test_committed_table_parses():
    t = autotune.read_table(autotune.TABLE)
    for k, (tile, split, stride) in t.items():
        assert k.count("/") == 6 and tile % 64 == 0 and 64 <= tile <= 256
        assert split in (1, 2, 4, 8, 16) and stride in (1, 8)
        assert (tile // 64) * split <= 16
