"""Shape selection of the autotuner from a cached table (no GPU)."""
import json

import pytest

from fd3d_tsn_amd import autotune as A
from fd3d_tsn_amd import cases as C


def test_key_round_trip():
    for shape in ((False, 16, 32), (True, 8, 64), (True, 4, 1)):
        assert A.parse(A.key(*shape)) == shape


def test_candidates_cover_wide_tiles():
    cs = A.candidates(C.c2_large_sw())
    assert (True, 8, 32) in cs and (False, 16, 32) in cs
    assert all(tz in A.WIDE_TZS for w, tz, _ in cs if w)


@pytest.fixture
def cache(tmp_path, monkeypatch):
    path = tmp_path / "cache.json"
    monkeypatch.setattr(A, "CACHE", str(path))
    return path


def test_choose_sets_width_bits(cache):
    case = C.c2_large_sw()
    tab = {"vel": {"16x32": 3.0, "w8x32": 2.5, "8x64": 3.2},
           "str": {"16x32": 4.0, "w8x16": 4.1, "4x8": 5.0},
           "pml": {"flat": 3.0, "zmarch": 3.5}}
    json.dump({"entries": {A.size_class(case): tab}}, open(cache, "w"))
    assert A.choose(case) == (1 | 8, 8, 32, 16, 32)
    # a fixed tile height (the batch engine's) uses 64-wide shapes only
    assert A.choose(case, tz=16) == (1, 16, 32, 16, 32)
    assert A.choose(case, tz=8) is None   # no 64 x 8 stress shape measured


def test_batch_plan_minimises_summed_time(cache):
    a, b = C.c2_large_sw(), C.c3_fvw()
    json.dump({"entries": {
        A.size_class(a): {"vel": {"16x32": 3.0, "8x32": 3.5}, "str": {"16x32": 4.0, "8x32": 4.2}, "pml": {}},
        A.size_class(b): {"vel": {"16x32": 1.0, "8x32": 0.2}, "str": {"16x32": 1.0, "8x32": 0.3}, "pml": {}},
    }}, open(cache, "w"))
    tz, plan = A.batch_plan([a, b], fused=False)
    assert tz == 8                      # 3.5 + 4.2 + 0.2 + 0.3 < 3 + 4 + 1 + 1
    assert plan[0] == (1, 8, 32, 8, 32) and plan[1] == (1, 8, 32, 8, 32)
    # with the fused sweep: both members (b has a free surface) ask for it,
    # in chunks of at most 128 rows
    tz, plan = A.batch_plan([a, b])
    assert plan[0] == (1 | 4, 8, 128, 8, 32) and plan[1] == (1 | 4, 8, 128, 8, 32)


def test_choose_takes_fused_sweep_when_faster(cache):
    case = C.c2_large_sw()   # full space
    tab = {"vel": {"16x32": 3.0, "w8x32": 2.5}, "str": {"16x32": 4.0, "w8x16": 4.1},
           "pml": {"flat": 3.0}, "fused": {"64": 5.0, "256": 4.8}}
    json.dump({"entries": {A.size_class(case): tab}}, open(cache, "w"))
    # 4.8 < 0.95 (2.5 + 4.0): fused (bit 2), rows per block from the grid
    # (752 fused tiles, 510 rows: one chunk, 3 waves x 514 rows)
    assert A.fused_ny(case) == 510
    assert A.choose(case) == (1 | 8 | 4, 8, 510, 16, 32)
    # the batch engine runs separate sweeps only
    assert A.choose(case, tz=16) == (1, 16, 32, 16, 32)
    tab["fused"] = {"64": 6.3}          # within the 5% margin of 6.5: separate
    json.dump({"entries": {A.size_class(case): tab}}, open(cache, "w"))
    assert A.choose(case) == (1 | 8, 8, 32, 16, 32)


def test_fused_ny_minimises_waves_times_rows():
    # 512x256x256: 192 tiles, 254 rows.  Measured fused sweep: 64 rows 0.667 ms,
    # 32 rows 0.693, 43 rows 0.718, 256 rows 0.836, 128 rows 0.848 - the order
    # of waves x (rows + 4): 204, 216, 235, 258, 264
    c = C.c3_fvw()
    assert A.fused_ny(c) == 64
    c = C.c2_large_sw().replace(nxtT=248, nytT=274, nztT=310)   # 116 tiles, 272 rows
    assert A.fused_ny(c) == 136                                 # 2 chunks: 232 blocks, one wave


def test_fused_chosen_with_free_surface(cache):
    case = C.c3_fvw()            # free surface
    tab = {"vel": {"16x32": 3.0}, "str": {"16x32": 4.0}, "pml": {}, "fused": {"64": 1.0}}
    json.dump({"entries": {A.size_class(case): tab}}, open(cache, "w"))
    assert A.choose(case)[0] & 4
    # in a batch too, in chunks of at most 128 rows
    t = A.batch_member(case, A.choose(case, 16), 16)
    assert t[0] & 4 and t[2] == 128


def test_batch_member_fused_chunks():
    small = C.c1_plain_sw()            # full space, 62 interior rows
    assert A.batch_member(small, None, 16) == (1 | 4, 16, 62, 16, 32)
    assert A.batch_member(small, (1, 8, 16, 8, 32), 8, fused=False) == (1, 8, 16, 8, 32)
    assert A.batch_member(C.c3_fvw(), (1, 8, 16, 8, 32), 8) == (1 | 4, 8, 128, 8, 32)   # free surface


def test_tiled_pml_keeps_the_tile_width_bits(cache):
    """When the tiled PML sweep measured faster, choose() adds mode bit 1 to
    the interior shape it picked (128-wide tiles included)."""
    case = C.c2_large_sw()
    tab = {"vel": {"w8x32": 2.5, "16x32": 3.0}, "str": {"w8x16": 3.9, "16x32": 4.0},
           "pml": {"flat": 3.0, "tiled": 2.9}}
    json.dump({"entries": {A.size_class(case): tab}}, open(cache, "w"))
    assert A.choose(case) == (1 | 2 | 8 | 16, 8, 32, 8, 16)


def test_fused_strip_width_only_when_one_percent_faster(cache):
    """choose() keeps the library's strips of 4 tiles unless another width
    measured at least 1% faster for the class, and then sets its mode bits
    5-6 (2 tiles: 32, 8 tiles: 64) beside the fused sweep's bit 2."""
    case = C.c2_large_sw()
    base = {"vel": {"16x32": 3.0}, "str": {"16x32": 3.0}, "pml": {"flat": 3.0}, "fused": {"256": 4.6}}
    ny = A.fused_ny(case)
    for strips, bits in (({"4": 4.70, "2": 4.61, "8": 4.75}, 32),
                         ({"4": 0.600, "2": 0.620, "8": 0.570}, 64),
                         ({"4": 4.70, "2": 4.67, "8": 4.69}, 0),     # within 1%: strips of 4
                         (None, 0)):
        tab = dict(base)
        if strips:
            tab["fused_strip"] = strips
        json.dump({"entries": {A.size_class(case): tab}}, open(cache, "w"))
        assert A.choose(case) == (1 | 4 | bits, 16, ny, 16, 32), strips
    # the batch engine's tile height: no fused sweep, no strip bits
    assert A.choose(case, 16)[0] & 100 == 0
