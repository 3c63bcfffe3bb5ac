"""Autotuner of the launch shapes of the 3D sweeps, per problem size class.

The interior sweeps (tiles.h) take a tile of 64 x tz (tz = 4, 8, 16) or
128 x tz (tz = 4, 8) cells of the x-z plane and a number ny of rows each
block marches along y; the
PML sweeps run one thread per cell or march z in register windows.  The
best shape depends on the grid: narrow grids waste lanes on tall tiles,
short y extents leave too few blocks when ny is large.  Results never
depend on the shape (the kernels are bit-identical across shapes), so a
problem can be tuned live, in a few steps, by timing every candidate
with HIP events on its own stream.

Measured tables are cached per device and size class in a JSON file
(``tuning_cache.json`` next to this module, or $FD3D_TUNING_CACHE).  A
size class rounds each interior extent to the nearest of a fixed ladder
of sizes and keeps the PML width and the free-surface setting.

The fused velocity + stress sweep (mode bit 2) is timed too, at several rows per block; :func:`choose` takes it when it
beats the best pair of separate sweeps by 5%, with rows per block from
the problem's own geometry (:func:`fused_ny`), and in strips of the
width measured fastest for the class (:func:`tune_strips`, mode bits 5-6).

* :func:`tune` measures one problem and stores its table;
* :func:`tune_strips` times the fused sweep's strip widths on a problem
  whose class takes the fused sweep and adds them to its table;
* :func:`choose` returns the best (tz, ny) of a problem's class, or the
  best ny for a given tz (the batch engine uses one tz for all members);
* :func:`apply` sets a solver's shape from the cache (or leaves the
  library's default when the class was never measured);
* :func:`batch_plan` picks the common tz of a batch that minimises the
  summed sweep time of its members, and each member's ny for that tz.
"""
from __future__ import annotations

import json
import os
from typing import Dict, Iterable, List, Optional, Tuple

from . import capi
from . import cases as C

HERE = os.path.dirname(os.path.abspath(__file__))
CACHE = os.environ.get("FD3D_TUNING_CACHE", os.path.join(HERE, "tuning_cache.json"))
LADDER = (32, 48, 64, 96, 128, 160, 192, 256, 320, 384, 512, 640, 768, 1024, 1536, 2048)
TZS = (4, 8, 16)
NYS = (8, 16, 32, 64, 128)
FUSED_NYS = (32, 64, 128, 256, 512)


def size_class(case: C.Case) -> str:
    def near(n):
        return min(LADDER, key=lambda v: abs(v - n) / v)
    return f"{near(case.nxtT)}x{near(case.nytT)}x{near(case.nztT)}/pml{case.nabc}/{'fs' if case.fspace else 'free'}"


def _load() -> Dict:
    if os.path.exists(CACHE):
        with open(CACHE) as f:
            return json.load(f)
    return {"entries": {}}


def _save(db: Dict):
    tmp = CACHE + ".tmp"
    with open(tmp, "w") as f:
        json.dump(db, f, indent=1, sort_keys=True)
    os.replace(tmp, CACHE)


WIDE_TZS = (4, 8)


def key(wide: bool, tz: int, ny: int) -> str:
    """Table key of a shape: "16x32" is 64 x 16 tiles marching 32 rows,
    "w8x32" is 128 x 8 tiles marching 32 rows."""
    return f"{'w' if wide else ''}{tz}x{ny}"


def parse(k: str) -> Tuple[bool, int, int]:
    wide = k.startswith("w")
    tz, ny = (int(v) for v in k.lstrip("w").split("x"))
    return wide, tz, ny


def candidates(case: C.Case) -> List[Tuple[bool, int, int]]:
    nyi = case.nyt - 2 - case.nabc
    out = []
    for wide, tzs in ((False, TZS), (True, WIDE_TZS)):
        for tz in tzs:
            for ny in NYS:
                if ny <= max(8, 2 * nyi):
                    out.append((wide, tz, ny))
    return out


def tune(solver: "capi.Solver", steps: int = 3, warmup: int = 1, save: bool = True) -> Dict:
    """Time every candidate shape on a live problem (its state advances by
    the steps taken; results do not depend on the shape).  Returns and
    caches {"vel": {"tz x ny": ms}, "str": {...}, "pml": {"flat": ms, "tiled": ms}}
    and "fused": {"ny": ms} (the fused sweep, plus the top-plane stresses
    under a free surface)."""
    case = solver.case
    before = solver.get_tuning()
    table = {"vel": {}, "str": {}, "pml": {}}
    nyi = case.nyt - 2 - case.nabc
    table["fused"] = {}
    for ny in FUSED_NYS:
        if ny > max(32, 2 * nyi):
            continue
        solver.set_tuning(5, 16, ny, 16, ny)
        solver.step(warmup, outputs=False)
        solver.set_profiling(True)
        solver.step(steps, outputs=False)
        kt = solver.kernel_times()
        solver.set_profiling(False)
        ms = kt["fused_interior"][0] / kt["fused_interior"][1]
        if "str_top" in kt:          # free surface: the top planes' stresses
            ms += kt["str_top"][0] / kt["str_top"][1]
        table["fused"][str(ny)] = ms
    for wide, tz, ny in candidates(case):
        for pml_mode in ((1, 3) if (wide, tz, ny) == (False, 16, 32) else (1,)):
            solver.set_tuning(pml_mode | (24 if wide else 0), tz, ny, tz, ny)
            solver.step(warmup, outputs=False)
            solver.set_profiling(True)
            solver.step(steps, outputs=False)
            kt = solver.kernel_times()
            solver.set_profiling(False)
            if pml_mode == 1:
                table["vel"][key(wide, tz, ny)] = kt["vel_interior"][0] / kt["vel_interior"][1]
                table["str"][key(wide, tz, ny)] = kt["str_interior"][0] / kt["str_interior"][1]
            if "pml_vel" in kt:
                pm = (kt["pml_vel"][0] / kt["pml_vel"][1]) + (kt["pml_str"][0] / kt["pml_str"][1])
                table["pml"]["tiled" if pml_mode & 2 else "flat"] = pm
    solver.set_tuning(*before)
    if save:
        db = _load()
        db.setdefault("entries", {})[size_class(case)] = table
        _save(db)
    return table


# Strip widths of the fused sweep's tile order and their mode bits 5-6
# (fused.h strip_of_mode): tiles are walked in z within strips this many
# x-tiles wide.  The best width depends on how the tile count splits over
# the 8 XCDs: 1024x512x512 (16 x 47 tiles, 94 per XCD) ran 4.61 ms in
# strips of 2 against 4.70 ms in strips of 4, 512x256x256 (8 x 24 tiles
# in 4 y chunks, 96 per XCD) 0.656 ms in strips of 4 against 0.716 ms in
# strips of 2.
STRIPS = {4: 0, 2: 1 << 5, 8: 2 << 5}


STRIP_MARGIN = 0.99


def tune_strips(solver: "capi.Solver", steps: int = 5, warmup: int = 2, rounds: int = 3,
                save: bool = True) -> Optional[Dict]:
    """Time the fused sweep in each strip width of STRIPS at the shape
    :func:`choose` gives the solver's class (a class that does not take
    the fused sweep is left alone): `rounds` passes over the widths, the
    best average launch of each, stored as the entry's "fused_strip"
    {"width": ms}.  Results do not depend on the width."""
    case = solver.case
    t = choose(case)
    if t is None or not t[0] & 4:
        return None
    before = solver.get_tuning()
    res = {}
    solver.step(warmup, outputs=False)
    for _ in range(rounds):
        for w, bits in STRIPS.items():
            solver.set_tuning((t[0] & ~96) | bits, *t[1:])
            solver.step(1, outputs=False)
            solver.set_profiling(True)
            solver.step(steps, outputs=False)
            kt = solver.kernel_times()
            solver.set_profiling(False)
            ms = kt["fused_interior"][0] / kt["fused_interior"][1]
            res[str(w)] = min(ms, res.get(str(w), ms))
    solver.set_tuning(*before)
    if save:
        db = _load()
        db.setdefault("entries", {}).setdefault(size_class(case), {})["fused_strip"] = res
        _save(db)
    return res


def _best(tab: Dict[str, float], tz: Optional[int] = None) -> Optional[Tuple[bool, int, int, float]]:
    """Fastest (wide, tz, ny, ms) of a table; with tz given, among the
    64-wide shapes of that height only (the batch engine's tiles)."""
    best = None
    for k, ms in tab.items():
        w, t, n = parse(k)
        if tz is not None and (w or t != tz):
            continue
        if best is None or ms < best[3]:
            best = (w, t, n, ms)
    return best


def lookup(case: C.Case) -> Optional[Dict]:
    return _load().get("entries", {}).get(size_class(case))


FUSED_MARGIN = 0.95


def fused_ny(case: C.Case, ncu: int = 256) -> int:
    """Rows per block of the fused sweep from the problem's own geometry.
    The sweep holds one workgroup per CU for the whole of its y chunk, and
    each chunk re-reads a 4-row ring, so its time goes as waves x (rows +
    4) with waves = ceil(blocks / CUs); the row count minimising that is
    taken (the same rule as the library's default).  A row count measured
    on another member of the size class does not carry over: a 248x274x310
    grid at the class's 128 rows made 348 blocks, a second wave a third
    full, and ran 16% slower than the separate sweeps.  The model orders
    the measured row counts of the 1024x512x512 and 512x256x256 grids as
    measured, to within 0.3%."""
    tiles = max(1, -(-case.nxtT // 64) * -(-case.nztT // 11))
    nyi = max(1, case.nyt - 2 - case.nabc)
    best = None
    for fgy in range(1, 33):
        rows = -(-nyi // fgy)
        chunks = -(-nyi // rows)
        cost = -(-(tiles * chunks) // ncu) * (rows + 4)
        if best is None or cost < best[0]:
            best = (cost, rows)
    return best[1]


def choose(case: C.Case, tz: Optional[int] = None) -> Optional[Tuple[int, int, int, int, int]]:
    """(mode, tz_vel, ny_vel, tz_str, ny_str) from the cache, or None."""
    e = lookup(case)
    if not e or not e["vel"]:
        return None
    wv, tv, nv, _ = _best(e["vel"], tz) or (None, None, None, None)
    ws, ts, ns, _ = _best(e["str"], tz) or (None, None, None, None)
    if tv is None or ts is None:
        return None
    mode = 1 | (8 if wv else 0) | (16 if ws else 0)
    if e.get("pml") and e["pml"].get("tiled", 1e30) < e["pml"].get("flat", 1e30):
        mode |= 2   # tiled PML sweeps; the tile-width bits stay
    # the fused sweep (single problems only: the batch engine runs the
    # separate sweeps) when it beats the best separate pair
    # (rows per block from the problem's geometry, fused_ny) when its best
    # measured time beats the best separate pair by FUSED_MARGIN
    fz = e.get("fused")
    if tz is None and fz:
        fms = min(fz.values())
        if fms < FUSED_MARGIN * (_best(e["vel"])[3] + _best(e["str"])[3]):
            # another strip width than the library's 4 only when it
            # measured at least 1% faster
            fs = e.get("fused_strip")
            strip = 0
            if fs and "4" in fs:
                w = min(fs, key=fs.get)
                if fs[w] < STRIP_MARGIN * fs["4"]:
                    strip = STRIPS[int(w)]
            return (mode | 4 | strip, tv, fused_ny(case), ts, ns)
    return (mode, tv, nv, ts, ns)


def apply(solver: "capi.Solver", tz: Optional[int] = None) -> bool:
    t = choose(solver.case, tz)
    if t is None:
        return False
    solver.set_tuning(*t)
    return True


FUSED_BATCH_NY = 128


def batch_member(case: C.Case, t: Optional[Tuple[int, int, int, int, int]], tz: int,
                 fused: bool = True) -> Optional[Tuple[int, int, int, int, int]]:
    """A member's shape in a batch: t (from choose at the batch's tz) and,
    when fused is set, the fused interior sweep (mode bit 2) in y chunks of
    at most FUSED_BATCH_NY rows (in ny_vel).
    In a batch the fused tiles of all members fill the GPU together, so it
    pays even for grids too small to fill it alone; chunks of a bounded
    length keep members of different lengths balanced.  When the batch
    holds too little fused work (fd3d_batch_set_fused_items) it runs the
    separate sweeps, both with the member's ny_str rows."""
    if not fused:
        return t
    nyi = case.nyt - 2 - case.nabc
    if nyi <= 0:
        return t
    base = t if t is not None else (1, tz, 32, tz, 32)
    return (base[0] | 4, base[1], min(nyi, FUSED_BATCH_NY), base[3], base[4])


def batch_plan(cases: Iterable[C.Case], fused: bool = True) -> Tuple[int, List[Optional[Tuple[int, int, int, int, int]]]]:
    """Common tile height for a batch and each member's shape at that height:
    the height with the least summed sweep time over the measured members,
    among the heights measured for every one of them (16 when none is);
    members run the fused sweep when fused is set (batch_member)."""
    cases = list(cases)
    tables = [e for e in (lookup(c) for c in cases) if e]
    totals = {}
    for tz in TZS:
        tot = 0.0
        for e in tables:
            bv, bs = _best(e["vel"], tz), _best(e["str"], tz)
            if not (bv and bs):
                tot = None
                break
            tot += bv[3] + bs[3]
        if tot is not None and tables:
            totals[tz] = tot
    tz = min(totals, key=totals.get) if totals else 16
    return tz, [batch_member(c, choose(c, tz), tz, fused) for c in cases]


def main():
    """python -m fd3d_tsn_amd.autotune [batched|served N|single|large] [N] [strips] [new]:
    tune the size classes of a workload on the GPU and update the cache
    (with "strips", only the fused sweep's strip widths of the classes that
    take it; with "new", only classes the cache does not hold yet, or
    whose strip widths it does not hold)."""
    import sys
    import time
    which = sys.argv[1] if len(sys.argv) > 1 else "batched"
    n = int(sys.argv[2]) if len(sys.argv) > 2 else 48
    if which == "batched":
        cs = C.c4_batched()
    elif which == "served":
        from . import serve
        cs = [r.case for r in serve.c5_requests(2000)[:n]]
    elif which == "single":
        cs = [C.c1_plain_sw(), C.c2_large_sw(), C.c3_fvw()]
    else:
        cs = [C.c2_large_sw()]
    flags = set(sys.argv[3:])
    strips_only = "strips" in flags
    only_new = "new" in flags
    seen = set()
    for c in cs:
        k = size_class(c)
        if k in seen:
            continue
        seen.add(k)
        e = lookup(c)
        if only_new and e is not None and (not strips_only or "fused_strip" in e):
            continue
        t0 = time.perf_counter()
        if strips_only:
            # the fused sweep's strip widths only, for classes already in
            # the cache that take it
            with capi.Solver(c.replace(ntfd=16, stations=[])) as s:
                res = tune_strips(s)
            if res:
                print(f"[autotune] {k}: fused strips " + " ".join(f"{w}:{ms:.3f}" for w, ms in res.items()) +
                      f" ({time.perf_counter()-t0:.1f} s)", flush=True)
            continue
        with capi.Solver(c.replace(ntfd=8, stations=[])) as s:
            tab = tune(s)
        bv, bs = _best(tab["vel"]), _best(tab["str"])
        fz = tab.get("fused") or {}
        fus = (" fused " + " ".join(f"{ny}:{ms:.3f}" for ny, ms in fz.items())) if fz else ""
        print(f"[autotune] {k}: vel {key(*bv[:3])} {bv[3]:.3f} ms, str {key(*bs[:3])} {bs[3]:.3f} ms{fus} "
              f"({time.perf_counter()-t0:.1f} s)", flush=True)


if __name__ == "__main__":
    main()
