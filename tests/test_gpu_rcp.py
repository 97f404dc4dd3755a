"""The setup stage's 1/x (axr::rcp_rn: hardware reciprocal + one FMA Newton
step inside the exponent range where that is exact, the division elsewhere)
gives the bits of the IEEE division for all 2^32 inputs on this GPU; the
coverage, depth and winner parity of every other test rests on it."""
import subprocess

import pytest

from axiomr_amd import build as B

pytestmark = pytest.mark.gpu


def test_rcp_rn_is_the_division_for_every_float():
    exe = B.RCP_BIN
    assert exe.exists(), "build the library first"
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=100)
    print(r.stdout, flush=True)
    assert r.returncode == 0, r.stdout + r.stderr
    tested, bad = r.stdout.split()[1], r.stdout.split()[3]
    assert int(tested) == 1 << 32 and int(bad) == 0
