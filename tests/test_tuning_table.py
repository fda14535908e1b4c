"""The shipped schedule table (tuning_gfx950.txt beside libnbh.so) is one the
library accepts: 'Nmax grp_lds mid_max cmin cmax' per line (the same after
'single' for products run alone), Nmax increasing in each table, every knob a
power of two, cmin <= cmax, tiles within a gfx950 block's LDS (160 KiB: 8 B
per point plus one pad double per 16)."""
import os

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
TABLE = os.path.join(ROOT, "nativebigint.jl_amd", "tuning_gfx950.txt")
LDS = 160 * 1024


def _pow2(x):
    return x > 0 and x & (x - 1) == 0


def test_table_is_well_formed():
    if not os.path.exists(TABLE):
        pytest.skip("no table shipped")
    rows, single = [], []
    for line in open(TABLE):
        if not line.strip() or line.startswith("#"):
            continue
        w = line.split()
        dst = single if w[0] == "single" else rows
        v = [int(x) for x in (w[1:] if w[0] == "single" else w)]
        assert len(v) == 5, line
        dst.append(v)
    assert rows
    for tab in (rows, single):
        assert [r[0] for r in tab] == sorted({r[0] for r in tab})
    for nmax, grp, mid, cmin, cmax in rows + single:
        assert all(_pow2(x) for x in (grp, mid, cmin, cmax)), (nmax, grp, mid, cmin, cmax)
        assert cmin <= cmax
        for pts in (grp, mid):
            assert 8 * (pts + pts // 16 + 1) <= LDS
        assert mid <= 16 * 1024          # one mid thread per 16 points, at most 1024 threads
