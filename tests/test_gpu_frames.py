"""The frame loop of the reference application (src/main.cpp:202-221,
28-29) as the `mpm_run` driver over the C++ mirror header: BGEO frames
`FrameNNNNNN.bgeo` written whenever int(t / (1/fps)) reaches the next frame
number, at the fixed dt and under the CFL step-size controller (the cadence
follows simulated time, not steps).  Frames are read back with the BGEO
layout of the host test (tests/test_capi_host.py)."""
import os
import re
import subprocess

import numpy as np
import pytest

from tests.test_capi_host import read_bgeo

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
RUN = os.path.join(ROOT, "mpm_cuda_amd", "mpm_run")


@pytest.mark.parametrize("dt_mode", ["fixed", "cfl"])
def test_frame_cadence_and_contents(gpu, tmp_path, dt_mode):
    out = tmp_path / "Animations"
    p = subprocess.run([RUN, "--scene", "jello", "--grid", "64", "--time", "0.1", "--fps", "36", "--out", str(out),
                        "--dt-mode", dt_mode], capture_output=True, text=True, timeout=300)
    print(p.stdout, p.stderr)
    assert p.returncode == 0, p.stderr
    n = int(re.search(r"particles: (\d+)", p.stdout).group(1))
    steps = int(re.search(r"steps: (\d+)", p.stdout).group(1))
    assert 25000 < n < 40000                      # 0.25^3 box, 8 particles per cell at h = 1/64
    # frames 0..3: int(t * 36) reaches 3 before t passes 0.1, not 4
    files = sorted(os.listdir(out))
    assert files == [f"Frame{k:06d}.bgeo" for k in range(4)], files
    if dt_mode == "fixed":
        assert steps in (1000, 1001)                # t += 1e-4 until t > 0.1 (rounding of the sum)
    else:
        assert steps >= 1000                        # dt <= max_dt
    ys = []
    for f in files:
        magic, _, _, counts, pts, _ = read_bgeo(str(out / f))
        assert counts[0] == n and np.isfinite(pts).all()
        assert np.all(pts[:, 3] == 1.0)             # homogeneous w
        assert pts[:, :3].min() > 0.0 and pts[:, :3].max() < 1.0
        ys.append(float(pts[:, 1].mean()))
    # the box starts at y in [0.1, 0.35] moving down at 2 m/s, lands on the
    # sticky floor band and rebounds: every frame shows it somewhere else
    assert 0.2 < ys[0] < 0.23
    assert min(abs(a - b) for a, b in zip(ys, ys[1:])) > 1e-3, ys
