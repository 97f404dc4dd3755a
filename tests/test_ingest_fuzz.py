"""Malformed inputs through the file loaders (axiomr_amd/csrc/image.cpp,
jpeg.cpp, ingest.cpp) built with AddressSanitizer and UBSan: truncated,
byte-flipped and length-corrupted PNG, JPEG and PPM files, and OBJ/MTL text
with bad indices, missing fields and garbage.  Each must load or be
rejected with an error; none may touch memory it does not own."""
import subprocess
from pathlib import Path

import numpy as np

from axiomr_amd import scenes as S
from axiomr_amd.ingest import write_png

HERE = Path(__file__).resolve().parent
CSRC = HERE.parent / "axiomr_amd" / "csrc"


def mutate(data: bytes, rng, n=25):
    out = [data[:k] for k in sorted({1, 8, 16, 33, len(data) // 4, len(data) // 2, len(data) - 1}) if 0 < k < len(data)]
    for _ in range(n):
        b = bytearray(data)
        for _ in range(int(rng.integers(1, 5))):
            b[int(rng.integers(0, len(b)))] = int(rng.integers(0, 256))
        out.append(bytes(b))
    return out


def test_loaders_under_sanitizers(tmp_path):
    exe = tmp_path / "ingest_fuzz"
    cmd = ["g++", "-std=c++17", "-O1", "-g", "-fsanitize=address,undefined", "-fno-sanitize-recover=undefined",
           f"-I{HERE.parent / 'include'}", f"-I{CSRC}", str(CSRC / "image.cpp"), str(CSRC / "jpeg.cpp"),
           str(CSRC / "ingest.cpp"), str(HERE / "native" / "ingest_fuzz.cpp"), "-lz", "-o", str(exe)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    rng = np.random.default_rng(7)
    files = []

    def add(name, data):
        p = tmp_path / name
        p.write_bytes(data)
        files.append(str(p))

    # images: PNG in several colour types, PPM, the JPEG fixtures
    srcs = []
    for k, img in enumerate([S.noise_texture(13, 9, 1), S.checker(16, 16, 4)]):
        p = tmp_path / f"src{k}.png"
        write_png(str(p), img)
        srcs.append(("png", p.read_bytes()))
    srcs.append(("ppm", b"P6\n4 3\n255\n" + bytes(range(36))))
    for j in sorted((HERE / "golden" / "jpeg").glob("*.jpg"))[:6]:
        srcs.append(("jpg", j.read_bytes()))
    for k, (ext, data) in enumerate(srcs):
        for m, v in enumerate(mutate(data, rng)):
            add(f"img{k}_{m}.{ext}", v)
    # OBJ text
    good = ("v 0 0 0\nv 1 0 0\nv 0 1 0\nv 1 1 0\nvt 0 0\nvt 1 0\nvt 0 1\nvn 0 0 1\n"
            "usemtl a\nf 1/1/1 2/2/1 3/3/1\nf 2/2/1 4/1/1 3/3/1\nf 1 2 3 4\n").encode()
    (tmp_path / "objmtl.mtl").write_text("newmtl a\nNs 2\nmap_Kd nothere.png\n")
    bad_lines = [b"f 1/1/1 9/9/9 3/3/3\n", b"f -1 -2 -3\n", b"f 0 0 0\n", b"f 1//1 2//1\n", b"v 1e40 nan inf\n",
                 b"f a/b/c d e f\n", b"usemtl\nf 1 2 3\n", b"vt\nvn\nf 1/2/3\n", b"f " + b"1 " * 5000 + b"\n",
                 b"\x00\xff\xfe garbage\n", b"f 1/1/1 2/2/1 3/3/1 2147483647/1/1\n"]
    for k, line in enumerate(bad_lines):
        add(f"objbad{k}.obj", good + line)
    for m, v in enumerate(mutate(good, rng, 30)):
        add(f"objmut{m}.obj", v)
    r = subprocess.run([str(exe), *files], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-4000:]
    assert "ERROR: AddressSanitizer" not in r.stderr and "runtime error" not in r.stderr, r.stderr[-4000:]
    out = r.stdout.split("\n")
    assert sum(ln.startswith("ok image") for ln in out) > 0 and sum(ln.startswith("ok obj") for ln in out) > 0
    assert sum(ln.startswith("err") for ln in out) > 0
