"""Importer of hgemm tune logs into the tuning database.

The reference's `tune hgemm` prints one line per candidate (hgemm.cu:300-322):
`[test] <name>`, `[failed] <name>`, `[result] <name>: <G> GFLOP/s` and finally
`[best] <name>: <G> GFLOP/s`, where <name> is `BM=.. BN=.. BK=.. WM=.. WN=..
stages=.. group_m=..` (hgemm_better_reuse.cu:378-454); its logs for square sizes
are kept as routines/hgemm/tunings/<size>.txt, some with lines run together.
This library's `libthom tune hgemm` prints the same line kinds with names
`BM=.. BN=.. BK=.. split_k=.. stages=..`.  Both parse here.

import_tune_log() turns every [result] into a row of the CLI's tuning
database (`M,N,K,bm,bn,bk,split_k,stages,ms`, read by `libthom test hgemm
tuned` from $LTHIP_TUNING_DIR/hgemm.csv): ms = 2 M N K / GFLOP/s.  The warp
tiling (WM, WN) and block swizzle (group_m) of the reference have no
counterpart here and are dropped; rows whose (BM, BN, BK, stages) the library
does not build can be left out.  The timings keep whatever hardware the log
was measured on (the reference's logs: an NVIDIA GPU), so an imported file is
for comparison or as a seed, not a measurement of this device.

Usage: python -m libthom_amd.tunelog LOG... --out DB.csv [--size M N K] [--library-only]
"""
from __future__ import annotations

import argparse
import os
import re
from typing import Dict, Iterable, List, Optional, Sequence, Tuple

DB_HEADER = "M,N,K,bm,bn,bk,split_k,stages,ms"
_LINE = re.compile(r"\[(result|best)\]\s*([^:\[\]]+?)\s*:\s*([0-9]+(?:\.[0-9]*)?)\s*GFLOP/s")
_KV = re.compile(r"([A-Za-z_]+)=(-?[0-9]+)")


def parse_name(name: str) -> Dict[str, int]:
    """`BM=64 BN=64 BK=32 WM=32 WN=32 stages=2 group_m=4` -> {"BM": 64, ...}."""
    return {k: int(v) for k, v in _KV.findall(name)}


def parse_tune_log(text: str) -> Tuple[List[Tuple[Dict[str, int], float]], Optional[Tuple[Dict[str, int], float]]]:
    """([(params, GFLOP/s) for every [result]], the [best] entry or None).
    Lines that were written without a newline between them are split apart."""
    results, best = [], None
    for kind, name, g in _LINE.findall(text):
        params = parse_name(name)
        if not {"BM", "BN", "BK"} <= params.keys():
            raise ValueError(f"tune log entry without BM/BN/BK: {name!r}")
        if kind == "result":
            results.append((params, float(g)))
        else:
            best = (params, float(g))
    return results, best


def size_from_path(path: str) -> Tuple[int, int, int]:
    """The reference names its logs by the square size: tunings/4096.txt."""
    stem = os.path.splitext(os.path.basename(path))[0]
    if not stem.isdigit():
        raise ValueError(f"{path}: give --size, the file name is not a square size")
    n = int(stem)
    return n, n, n


def library_tiles() -> set:
    """(bm, bn, bk, stages) of every hgemm tile the library builds."""
    import ctypes
    from ._lib import HgemmConfig, lib
    L = lib()
    out = set()
    for i in range(L.lthip_hgemm_config_count()):
        c = HgemmConfig()
        if L.lthip_hgemm_config_get(i, ctypes.byref(c)) == 0:
            out.add((c.bm, c.bn, c.bk, c.stages))
    return out


def to_rows(results: Iterable[Tuple[Dict[str, int], float]], M: int, N: int, K: int,
            tiles: Optional[set] = None) -> Tuple[List[Tuple], int]:
    """Database rows (M, N, K, bm, bn, bk, split_k, stages, ms) and the number
    of results left out because the library does not build their tile."""
    rows, skipped = [], 0
    flop = 2.0 * M * N * K
    for p, g in results:
        key = (p["BM"], p["BN"], p["BK"], p.get("stages", 2))
        if tiles is not None and key not in tiles:
            skipped += 1
            continue
        if g <= 0:
            raise ValueError(f"non-positive GFLOP/s in tune log: {p} {g}")
        rows.append((M, N, K, key[0], key[1], key[2], p.get("split_k", 1), key[3], flop / (g * 1e6)))
    return rows, skipped


def import_tune_log(path: str, out_csv: str, size: Optional[Sequence[int]] = None,
                    tiles: Optional[set] = None) -> dict:
    """Appends the log's [result] rows to out_csv (header written for a new
    file) and returns a summary; the log's [best] must be its fastest [result]."""
    with open(path) as f:
        results, best = parse_tune_log(f.read())
    M, N, K = size if size is not None else size_from_path(path)
    if best is not None and results:
        fastest = max(g for _, g in results)
        if abs(best[1] - fastest) > 1e-6 * fastest:
            raise ValueError(f"{path}: [best] {best[1]} is not the fastest [result] {fastest}")
    rows, skipped = to_rows(results, M, N, K, tiles)
    new = not os.path.exists(out_csv) or os.path.getsize(out_csv) == 0
    os.makedirs(os.path.dirname(os.path.abspath(out_csv)), exist_ok=True)
    with open(out_csv, "a") as f:
        if new:
            f.write(DB_HEADER + "\n")
        for r in rows:
            f.write(",".join(str(v) for v in r[:8]) + f",{r[8]:.6f}\n")
    return {"path": path, "size": (M, N, K), "results": len(results), "imported": len(rows), "skipped": skipped,
            "best": None if best is None else {"params": best[0], "GFLOP/s": best[1]}}


def main(argv=None):
    ap = argparse.ArgumentParser(description="import hgemm tune logs into the tuning database")
    ap.add_argument("logs", nargs="+")
    ap.add_argument("--out", required=True)
    ap.add_argument("--size", nargs=3, type=int)
    ap.add_argument("--library-only", action="store_true", help="keep only tiles this library builds")
    a = ap.parse_args(argv)
    tiles = library_tiles() if a.library_only else None
    for p in a.logs:
        print(import_tune_log(p, a.out, a.size, tiles))


if __name__ == "__main__":
    main()
