"""The tool ranges named after the reference's phases (AXR_ROCTX=1): a
process that has them on draws the same pixels as one that has them off,
in both raster modes.  The switch is read once
per process, so the ranged draw runs in a child process."""
import json
import os
import subprocess
import sys
from pathlib import Path

import numpy as np
import pytest

import harness as H
from axiomr_amd import scenes as S

pytestmark = pytest.mark.gpu
ROOT = Path(__file__).resolve().parents[1]

CHILD = r"""
import json, sys
sys.path.insert(0, {root!r}); sys.path.insert(0, {tests!r})
import numpy as np
import axiomr_amd as axr
import harness as H
from axiomr_amd import scenes as S
out = {{}}
ctx = axr.Context(0)
for mode in (0, 1):
    ctx.set_config(raster_mode=mode)
    c, d, i = H.render_gpu(ctx, S.icosphere_scene(n=24, width=320, height=240, tex=256))
    out[f"plain{{mode}}"] = [int(c.astype(np.uint64).sum()), int(d.view(np.uint32).astype(np.uint64).sum()),
                            int(i.astype(np.uint64).sum())]
ctx.close()
print(json.dumps(out))
"""


def _sums(ctx):
    out = {}
    for mode in (0, 1):
        ctx.set_config(raster_mode=mode)
        c, d, i = H.render_gpu(ctx, S.icosphere_scene(n=24, width=320, height=240, tex=256))
        out[f"plain{mode}"] = [int(c.astype(np.uint64).sum()), int(d.view(np.uint32).astype(np.uint64).sum()),
                              int(i.astype(np.uint64).sum())]
    return out


def test_phase_ranges_leave_pixels_unchanged():
    import axiomr_amd as axr
    print("phase ranges: child with AXR_ROCTX=1", flush=True)
    env = dict(os.environ, AXR_ROCTX="1")
    code = CHILD.format(root=str(ROOT), tests=str(ROOT / "tests"))
    r = subprocess.run([sys.executable, "-c", code], env=env, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr[-2000:]
    child = json.loads(r.stdout.strip().splitlines()[-1])
    ctx = axr.Context(0)
    try:
        assert child == _sums(ctx)
    finally:
        ctx.close()
