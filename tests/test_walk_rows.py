"""The coverage walk finds the row of a pixel index q inside a rectangle of
width w as trunc(fl(fl(q + 0.5) * fl(1 / w))) in float32 instead of an
integer division (raster.hip walk_span, the big-triangle loop of
raster_bin).  The claim that this equals q // w for every q <= 1024 and
1 <= w <= 1024 (the footprint limit) is checked here exhaustively with
float32 arithmetic, each operation rounded once as on the GPU."""
import numpy as np


def test_float_row_split_is_exact():
    q = np.arange(0, 1025, dtype=np.int64)
    for w in range(1, 1025):
        inv_w = np.float32(1.0) / np.float32(w)              # correctly rounded, as 1.0f / (float)w
        prod = (q.astype(np.float32) + np.float32(0.5)) * inv_w
        assert prod.dtype == np.float32
        dy = np.trunc(prod).astype(np.int64)
        assert np.array_equal(dy, q // w), w


def test_the_bound_is_needed():
    # a plain float division without the half-pixel offset would not do:
    # 3 * (1/3) rounds below 1 for some widths, so trunc gives the previous row
    bad = [w for w in range(1, 1025)
           if int(np.trunc(np.float32(w) * (np.float32(1.0) / np.float32(w)))) != 1]
    assert bad, "expected widths where q * fl(1/w) truncates below q // w"
