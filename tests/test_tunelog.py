"""Import of hgemm tune logs (the reference's routines/hgemm/tunings/<size>.txt
format, hgemm.cu:300-322, and this CLI's `tune hgemm` output) into the tuning
database the CLI's `tuned` implementation reads."""
import os

import pytest

from libthom_amd import tunelog

HERE = os.path.dirname(os.path.abspath(__file__))
EXCERPT = os.path.join(HERE, "golden", "hgemm_tune_4096_excerpt.txt")   # lines of the reference's 4096.txt


def test_parse_reference_excerpt():
    results, best = tunelog.parse_tune_log(open(EXCERPT).read())
    assert len(results) == 8
    assert results[0] == ({"BM": 64, "BN": 64, "BK": 64, "WM": 32, "WN": 32, "stages": 2, "group_m": 4}, 132630.20)
    assert best == ({"BM": 128, "BN": 128, "BK": 32, "WM": 64, "WN": 64, "stages": 2, "group_m": 8}, 193405.17)
    assert best[1] == max(g for _, g in results)


def test_parse_lines_run_together_and_own_format():
    # the reference's 512.txt has a [best] and a [result] on one line; this
    # CLI's names carry split_k and no warp tiling; [test]/[failed]/[skip] are ignored
    text = ("[test] BM=64 BN=64 BK=32 WM=32 WN=32 stages=2 group_m=4\n"
            "[failed] BM=64 BN=64 BK=32 WM=32 WN=32 stages=3 group_m=4\n"
            "[best] BM=64 BN=64 BK=64 WM=32 WN=32 stages=4 group_m=8: 15404.41 GFLOP/s"
            "[result] BM=64 BN=64 BK=32 WM=32 WN=32 stages=2 group_m=4: 145173.83 GFLOP/s\n"
            "[skip] BM=256 BN=256 BK=64 split_k=8 stages=2 (fewer K-slabs than splits)\n"
            "[result] BM=256 BN=256 BK=32 split_k=2 stages=4: 1089000.50 GFLOP/s\n")
    results, best = tunelog.parse_tune_log(text)
    assert [g for _, g in results] == [145173.83, 1089000.50]
    assert results[1][0] == {"BM": 256, "BN": 256, "BK": 32, "split_k": 2, "stages": 4}
    assert best[1] == 15404.41


def test_import_rows_and_filter(tmp_path):
    out = tmp_path / "hgemm.csv"
    tiles = {(128, 128, 32, 2), (128, 128, 32, 4)}
    s = tunelog.import_tune_log(EXCERPT, str(out), size=(4096, 4096, 4096), tiles=tiles)
    assert s["results"] == 8 and s["imported"] == 4 and s["skipped"] == 4
    lines = out.read_text().splitlines()
    assert lines[0] == "M,N,K,bm,bn,bk,split_k,stages,ms"   # the CLI's tuning database (cli/tuning.cpp)
    rows = [ln.split(",") for ln in lines[1:]]
    assert [tuple(map(int, r[3:8])) for r in rows] == [(128, 128, 32, 1, 2)] * 2 + [(128, 128, 32, 1, 4)] * 2
    # ms = 2 M N K / GFLOP/s: 193405.17 GFLOP/s at 4096^3 is 0.710642 ms
    assert float(rows[1][8]) == pytest.approx(2 * 4096 ** 3 / 193405.17e6, rel=1e-6)
    # a second import appends without a second header
    tunelog.import_tune_log(EXCERPT, str(out), size=(4096, 4096, 4096))
    assert out.read_text().count("M,N,K") == 1 and len(out.read_text().splitlines()) == 1 + 4 + 8


def test_size_from_reference_file_name_and_best_check(tmp_path):
    p = tmp_path / "1024.txt"
    p.write_text("[result] BM=64 BN=64 BK=32 stages=2: 100.0 GFLOP/s\n[best] BM=64 BN=64 BK=32 stages=2: 100.0 GFLOP/s")
    s = tunelog.import_tune_log(str(p), str(tmp_path / "db.csv"))
    assert s["size"] == (1024, 1024, 1024) and s["imported"] == 1
    p.write_text("[result] BM=64 BN=64 BK=32 stages=2: 100.0 GFLOP/s\n[best] BM=64 BN=64 BK=32 stages=2: 90.0 GFLOP/s")
    with pytest.raises(ValueError):
        tunelog.import_tune_log(str(p), str(tmp_path / "db.csv"))
    with pytest.raises(ValueError):
        tunelog.size_from_path(str(tmp_path / "log.txt"))


def test_library_tiles_contain_reference_overlap():
    # the library's built tiles (no GPU needed) include some of the reference's candidates
    tiles = tunelog.library_tiles()
    assert (128, 128, 32, 4) in tiles and (64, 64, 64, 2) in tiles
