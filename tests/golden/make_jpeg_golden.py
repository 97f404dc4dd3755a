"""Writes tests/golden/jpeg/: small JPEG files (PIL encodes them) and, for
each, the RGBA8 pixels the reference's own decoder gives (stb_image v2.25,
compiled from the reference tree into oracle/_ref by oracle.build()), as
SHA-256 digests plus the first row in clear, in cases.json.

    python tests/golden/make_jpeg_golden.py

Run where the reference tree is present; the fixtures it writes are data
(encoded images and digests of decoded pixels), checked by tests/test_jpeg.py.
"""
import hashlib
import io
import json
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))
import oracle as O  # noqa: E402

OUT = Path(__file__).resolve().parent / "jpeg"


def picture(w, h, seed):
    rng = np.random.default_rng(seed)
    y, x = np.mgrid[0:h, 0:w]
    base = np.stack([(x * 3 + y) % 256, (x * y // 7) % 256, (255 - x) % 256], -1)
    return np.clip(base + rng.integers(0, 60, (h, w, 3)), 0, 255).astype(np.uint8)


# name, size, PIL mode, save options
CASES = [
    ("base444", (37, 29), "RGB", dict(quality=90, subsampling=0)),
    ("base422", (37, 29), "RGB", dict(quality=75, subsampling=1)),
    ("base420", (37, 29), "RGB", dict(quality=50, subsampling=2)),
    ("prog420", (45, 31), "RGB", dict(quality=80, subsampling=2, progressive=True)),
    ("prog444", (45, 31), "RGB", dict(quality=95, subsampling=0, progressive=True)),
    ("gray", (30, 22), "L", dict(quality=85)),
    ("gray_prog", (30, 22), "L", dict(quality=85, progressive=True)),
    ("cmyk", (24, 20), "CMYK", dict(quality=90)),
    ("rgb_ids", (26, 18), "RGB", dict(quality=90, keep_rgb=True, subsampling=0)),
    ("restart", (64, 40), "RGB", dict(quality=70, subsampling=2, restart_marker_blocks=3)),
    ("restart_prog", (64, 40), "RGB", dict(quality=70, subsampling=2, progressive=True, restart_marker_rows=1)),
    ("q1", (40, 24), "RGB", dict(quality=1, subsampling=2)),
    ("dqt16", (40, 24), "RGB", dict(subsampling=2, qtables=[[min(1000, 300 + 7 * i) for i in range(64)],
                                                            [min(2000, 400 + 11 * i) for i in range(64)]])),
    ("q100", (40, 24), "RGB", dict(quality=100, subsampling=0)),
    ("optimized", (33, 33), "RGB", dict(quality=60, optimize=True)),
    ("tiny", (1, 1), "RGB", dict(quality=90)),
]


def main():
    from PIL import Image
    if O.stb_image() is None:
        O.build()
    OUT.mkdir(exist_ok=True)
    meta = []
    for i, (name, (w, h), mode, opts) in enumerate(CASES):
        im = Image.fromarray(picture(w, h, i), "RGB")
        if mode != "RGB":
            im = im.convert(mode)
        b = io.BytesIO()
        im.save(b, "JPEG", **opts)
        data = b.getvalue()
        (OUT / f"{name}.jpg").write_bytes(data)
        px = O.stbi_load_rgba(data)
        meta.append({"file": f"{name}.jpg", "width": int(px.shape[1]), "height": int(px.shape[0]),
                     "rgba_sha256": hashlib.sha256(px.tobytes()).hexdigest(),
                     "first_row": px[0].reshape(-1).tolist()[:64]})
    (OUT / "cases.json").write_text(json.dumps(meta, indent=1) + "\n")
    print(f"wrote {len(meta)} cases to {OUT}")


if __name__ == "__main__":
    main()
