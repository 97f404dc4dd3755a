"""Writes raster_kat.json: known-answer raster cases computed with exact
rational arithmetic (fractions.Fraction), independent of any float code.

The reference has no tests or golden images, so these fixtures come from
the mathematics of its raster rule (reference src/tiled_pipeline.cpp
:436-594): a pixel (px, py) belongs to a triangle when it lies in
[floor(minX), ceil(maxX)) x [floor(minY), ceil(maxY)) and all three edge
functions at its centre are >= 0 once the triangle is oriented to positive
area; zero-area triangles draw nothing; back faces (negative signed area,
y up) are culled; the nearest depth wins and exact ties go to the earlier
triangle.

Every corner is a multiple of 1/4 pixel on a 64 x 32 target and enters as
NDC with w = 1, so every float32 step of the setup and the edge functions is
exact and the float pipeline must reproduce these answers bit for bit.
Depths are constant per triangle and spaced 1/64 apart (exact duplicates
aside), far beyond the rounding of the interpolated depth.

    python tests/golden/make_golden.py
"""
from __future__ import annotations

import json
import math
from fractions import Fraction as Fr
from pathlib import Path

W, H = 64, 32
EMPTY = 0xFFFFFFFF


def q(v):
    return Fr(v).limit_denominator(4)


# (corners in pixels (x, y), constant depth); order = submission order
TRIS = [
    # a right triangle whose hypotenuse passes exactly through pixel centres
    ([(2.5, 2.5), (18.5, 2.5), (2.5, 18.5)], 0.50),
    # its mirror sharing the diagonal: centres on the shared edge belong to both
    ([(18.5, 2.5), (18.5, 18.5), (2.5, 18.5)], 0.45),
    # crosses the 16-pixel anchors at x = 32 and x = 48
    ([(28.0, 4.0), (52.25, 6.75), (40.0, 27.5)], 0.30),
    # exact duplicate submitted later with a nearer... same depth: loses every tie
    ([(28.0, 4.0), (52.25, 6.75), (40.0, 27.5)], 0.30),
    # a back face (clockwise): culled
    ([(5.0, 20.0), (5.0, 30.0), (20.0, 25.0)], 0.10),
    # nearer sliver over the first two
    ([(0.0, 9.75), (63.75, 10.25), (63.75, 11.0)], 0.20),
    # zero area (collinear): nothing
    ([(10.0, 25.0), (20.0, 26.0), (30.0, 27.0)], 0.05),
    # corners on pixel centres, one pixel wide
    ([(55.5, 20.5), (57.5, 20.5), (56.5, 22.5)], 0.15),
    # partly off the right edge (x up to 64, the last pixel column is 63)
    ([(60.0, 24.0), (64.0, 24.0), (64.0, 31.75)], 0.40),
    # depth tie between different triangles is avoided; deeper large one
    ([(0.25, 0.25), (63.75, 0.25), (32.0, 31.75)], 0.90),
]


def covers(tri, px, py):
    (x0, y0), (x1, y1), (x2, y2) = [(q(x), q(y)) for x, y in tri]
    xs, ys = (x0, x1, x2), (y0, y1, y2)
    if not (math.floor(min(xs)) <= px < math.ceil(max(xs)) and math.floor(min(ys)) <= py < math.ceil(max(ys))):
        return False
    area = (x1 * y2 - x2 * y1) + (x2 * y0 - x0 * y2) + (x0 * y1 - x1 * y0)
    if area == 0:
        return False
    s = 1 if area > 0 else -1
    cx, cy = Fr(px) + Fr(1, 2), Fr(py) + Fr(1, 2)
    e0 = (y1 - y2) * cx + (x2 - x1) * cy + (x1 * y2 - x2 * y1)
    e1 = (y2 - y0) * cx + (x0 - x2) * cy + (x2 * y0 - x0 * y2)
    e2 = (y0 - y1) * cx + (x1 - x0) * cy + (x0 * y1 - x1 * y0)
    return s * e0 >= 0 and s * e1 >= 0 and s * e2 >= 0


def backface(tri):
    (x0, y0), (x1, y1), (x2, y2) = [(q(x), q(y)) for x, y in tri]
    return (x1 - x0) * (y2 - y0) - (x2 - x0) * (y1 - y0) < 0


def main():
    ids = [[EMPTY] * W for _ in range(H)]
    depth = [[None] * W for _ in range(H)]
    for t, (tri, z) in enumerate(TRIS):
        if backface(tri):
            continue
        for py in range(H):
            for px in range(W):
                if covers(tri, px, py) and (depth[py][px] is None or z < depth[py][px]):
                    depth[py][px] = z
                    ids[py][px] = t * 8
    ndc = [[[2.0 * x / W - 1.0, 2.0 * y / H - 1.0, z] for x, y in tri] for tri, z in TRIS]
    out = {
        "description": "exact-arithmetic raster known answers; see make_golden.py",
        "width": W, "height": H,
        "triangles_ndc": ndc,
        "expected_ids": ids,
        "expected_depth": [[(d if d is not None else None) for d in row] for row in depth],
    }
    path = Path(__file__).with_name("raster_kat.json")
    path.write_text(json.dumps(out))
    covered = sum(v != EMPTY for row in ids for v in row)
    print(f"wrote {path} ({covered} covered pixels)")


if __name__ == "__main__":
    main()
