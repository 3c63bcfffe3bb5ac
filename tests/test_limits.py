"""Size limits and configuration checks of fd3d_create, through
fd3d_footprint (host-only: it validates the configuration and returns the
device bytes the problem would take, 0 when fd3d_create would refuse it).

Every 3D array is addressed with 32-bit byte offsets, so a padded array
must hold fewer than 2^30 floats; the PML split block of one region
likewise.  The reference itself has no such limit (it is bounded by host
memory); these tests pin where this build's limit falls."""
from fd3d_tsn_amd import capi
from fd3d_tsn_amd import cases as C


def fp(case):
    return capi.footprint(case)


def test_benchmark_grids_fit():
    b = fp(C.c2_large_sw())
    # two wavefield sets of 18 padded arrays (1152 x 532 x 552 floats each),
    # the PML splits (24 x 41.7M floats), the fault plane and the output ring
    assert 26e9 < b < 30e9, b
    assert fp(C.c3_fvw()) > 0 and fp(C.c1_plain_sw()) > 0


def test_padded_array_limit():
    # 1152 x 532 x 1064 floats = 0.61 * 2^30 per array: accepted
    assert fp(C.c2_large_sw().replace(nztT=1024, stations=[])) > 0
    # 2112 x 1044 x 552 floats = 1.14 * 2^30: refused
    assert fp(C.c2_large_sw().replace(nxtT=2048, nytT=1024, stations=[])) == 0


def test_invalid_configurations_refused():
    base = C.tiny()
    assert fp(base) > 0
    assert fp(base.replace(nabc=2)) == 0                         # PML too thin
    assert fp(base.replace(nxtT=1, nytT=1, nztT=1)) == 0         # no interior left
    assert fp(base.replace(stations=[(0, 1, 1)])) == 0           # station outside the grid
    assert fp(base.replace(stations=[(base.nxt, base.nyt, base.nzt)])) > 0
