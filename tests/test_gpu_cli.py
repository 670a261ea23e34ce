"""The athom tool (reference CLI, src/athom.cu:30-47) end to end on the GPU:
inputs written in the reference formats, statistics lines and the written
restart compared with the FP64 oracle started from the same file state."""
import os
import subprocess
import sys

import numpy as np
import pytest

from md_gpu_amd import fileio, systems as S

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ATHOM = os.path.join(ROOT, "md_gpu_amd", "bin", "athom")


def run_athom(name):
    print("[case] athom", name, file=sys.stderr, flush=True)
    r = subprocess.run([ATHOM, name], capture_output=True, text=True, timeout=240)
    assert r.returncode == 0, r.stderr
    return r.stdout


def stat_lines(out):
    rows = []
    for line in out.splitlines():
        parts = line.split("\t")
        if len(parts) == 3:
            rows.append((int(parts[0]), float(parts[1]), float(parts[2])))
    return rows


@pytest.mark.parametrize("kind", ["nacl_lf_berendsen", "water_vv_nh"])
def test_athom_matches_oracle(tmp_path, orc, kind):
    if kind == "nacl_lf_berendsen":
        s = S.nacl(nsites=12)
        timrun, timprn = 20, 5
    else:
        s = S.water(nmol=512)
        timrun, timprn = 20, 5
    name = str(tmp_path / kind)
    fileio.write_inputs(s, name, timrun=timrun, timprn=timprn)
    t = fileio.read_system(name)          # the exact state the tool reads (rounded by the format)
    out = run_athom(name)
    rows = stat_lines(out)
    assert [r[0] for r in rows] == list(range(0, timrun + 1, timprn))
    ref = orc.md_run(t, timrun + 1)       # t = 0..timrun inclusive (src/runGPU.cu:69)
    for step, T, ke in rows:
        assert T == pytest.approx(ref["temp"][step], rel=2e-4), step
        assert ke == pytest.approx(ref["ke"][step], rel=2e-4), step
    assert "force:" in out and "moment:" in out and "Total time" in out
    box, typ, q, pos, vel = fileio.read_cfg(name + ".xyz.cfg")
    d = pos.astype(np.float64) - ref["pos"]
    d -= box * np.rint(d / box)
    assert np.abs(d).max() < 5e-4
    assert np.array_equal(typ, t.type) and np.allclose(q, t.charge, atol=1e-6)


def test_athom_reports_bad_input(tmp_path):
    name = str(tmp_path / "none")
    r = subprocess.run([ATHOM, name], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "cannot open" in r.stderr
