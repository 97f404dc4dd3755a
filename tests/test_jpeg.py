"""JPEG textures (axiomr_amd/csrc/jpeg.cpp) against the reference's own
decoder, stb_image v2.25 (the reference's Texture calls stbi_load with 4
channels, src/texture.cpp:21-35):

- committed fixtures (tests/golden/jpeg/, written by make_jpeg_golden.py from
  the reference decoder compiled out of the reference tree): every pixel
  byte, by digest;
- where that build is present (oracle/_ref), a grid of encoder settings
  compared pixel for pixel;
- truncated and corrupted files through an AddressSanitizer/UBSan build of
  the decoder: an error or a well-formed image, never a memory error.
"""
import ctypes as C
import hashlib
import io
import json
import subprocess
from pathlib import Path

import numpy as np
import pytest

import oracle as O
from axiomr_amd import _native as N
from axiomr_amd.ingest import load_image

HERE = Path(__file__).resolve().parent
GOLDEN = HERE / "golden" / "jpeg"
CSRC = HERE.parent / "axiomr_amd" / "csrc"


def decode_bytes(tmp_path, data, name="x.jpg"):
    p = tmp_path / name
    p.write_bytes(data)
    return load_image(str(p))


@pytest.mark.parametrize("case", json.loads((GOLDEN / "cases.json").read_text()), ids=lambda c: c["file"])
def test_reference_decoder_fixtures(case):
    px = load_image(str(GOLDEN / case["file"]))
    assert px.shape == (case["height"], case["width"], 4)
    assert px[0].reshape(-1).tolist()[:64] == case["first_row"]
    assert hashlib.sha256(px.tobytes()).hexdigest() == case["rgba_sha256"]


def picture(w, h, seed):
    rng = np.random.default_rng(seed)
    y, x = np.mgrid[0:h, 0:w]
    base = np.stack([(x * 5 + 3 * y) % 256, (x * y // 5) % 256, (200 - 2 * x) % 256], -1)
    return np.clip(base + rng.integers(0, 80, (h, w, 3)), 0, 255).astype(np.uint8)


def test_grid_against_the_reference_decoder(tmp_path):
    Image = pytest.importorskip("PIL.Image")
    if O.stb_image() is None:
        pytest.skip("the reference decoder is built only where the reference tree is present")
    n = 0
    for size in [(1, 2), (9, 7), (16, 8), (17, 33), (100, 61), (257, 129)]:
        for mode in ["RGB", "L", "CMYK"]:
            for opts in [dict(quality=35), dict(quality=92, progressive=True), dict(quality=75, optimize=True),
                         dict(quality=80, restart_marker_blocks=2), dict(quality=97, progressive=True,
                                                                          restart_marker_rows=1)]:
                for sub in ([0, 1, 2] if mode != "L" else [None]):
                    im = Image.fromarray(picture(*size, n), "RGB")
                    if mode != "RGB":
                        im = im.convert(mode)
                    kw = dict(opts)
                    if sub is not None:
                        kw["subsampling"] = sub
                    b = io.BytesIO()
                    try:
                        im.save(b, "JPEG", **kw)
                    except OSError:   # a combination the encoder refuses
                        continue
                    want = O.stbi_load_rgba(b.getvalue())
                    got = decode_bytes(tmp_path, b.getvalue())
                    assert got.shape == want.shape and (got == want).all(), (size, mode, opts, sub)
                    n += 1
    assert n > 200


def test_rejects_what_is_not_a_decodable_jpeg(tmp_path):
    with pytest.raises(N.AxrError):
        decode_bytes(tmp_path, b"\xff\xd8\xff\xd9")                     # no frame
    lossless = bytes.fromhex("ffd8ffc3000b08000800080101110000ffd9")   # SOF3 (lossless)
    with pytest.raises(N.AxrError):
        decode_bytes(tmp_path, lossless)


def corrupt_variants(data, rng):
    out = [data[:k] for k in (3, 20, len(data) // 3, len(data) // 2, len(data) - 3)]
    for _ in range(40):
        b = bytearray(data)
        for _ in range(int(rng.integers(1, 6))):
            i = int(rng.integers(2, len(b)))
            b[i] = int(rng.integers(0, 256))
        out.append(bytes(b))
    # a segment length pointing past the end, a huge frame
    b = bytearray(data)
    i = b.find(b"\xff\xdb")
    b[i + 2:i + 4] = b"\xff\xf0"
    out.append(bytes(b))
    b = bytearray(data)
    i = max(b.find(b"\xff\xc0"), b.find(b"\xff\xc2"))
    b[i + 5:i + 9] = b"\xff\xff\xff\xff"
    out.append(bytes(b))
    return out


def test_corrupt_files_under_sanitizers(tmp_path):
    exe = tmp_path / "jpeg_fuzz"
    cmd = ["g++", "-std=c++17", "-O1", "-g", "-fsanitize=address,undefined", "-fno-sanitize-recover=undefined",
           f"-I{CSRC}", str(CSRC / "jpeg.cpp"), str(HERE / "native" / "jpeg_fuzz.cpp"), "-o", str(exe)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    rng = np.random.default_rng(3)
    files = []
    for src in sorted(GOLDEN.glob("*.jpg")):
        for k, v in enumerate(corrupt_variants(src.read_bytes(), rng)):
            p = tmp_path / f"{src.stem}_{k}.jpg"
            p.write_bytes(v)
            files.append(str(p))
    r = subprocess.run([str(exe), *files], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-4000:]
    assert "ERROR: AddressSanitizer" not in r.stderr and "runtime error" not in r.stderr, r.stderr[-4000:]
    lines = r.stdout.split("\n")
    assert sum(ln.startswith("ok") for ln in lines) > 0 and sum(ln.startswith("err") for ln in lines) > 0
