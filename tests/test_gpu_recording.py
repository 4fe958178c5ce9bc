"""The splinter driver on a keyframe recording in the reference's on-disk
format: a seeded synthetic scene exported with recording.write_recording is
rendered by splinter (reader, keyframe schedule, interpolated client camera,
hole retention across frames) and frame by frame through the Python
Renderer from the same files; the written PNGs must match byte for byte."""
import subprocess
from pathlib import Path

import numpy as np
import pytest
import torch

from imagebasedspatiotemporalinterpolation_amd import recording as RC
from imagebasedspatiotemporalinterpolation_amd import renderer as R
from imagebasedspatiotemporalinterpolation_amd import scene as S

pytestmark = pytest.mark.gpu

SPLINTER = Path(__file__).resolve().parent.parent / "imagebasedspatiotemporalinterpolation_amd" / "splinter"


@pytest.mark.parametrize("mode", [0, 3])
def test_splinter_on_a_recording_matches_the_python_renderer(tmp_path, mode):
    cfg = S.defaults(1, 21)
    cfg.width, cfg.height = 320, 176
    rec = tmp_path / "rec"
    RC.write_recording(rec, cfg, 3)
    info, _ = RC.read_camera(rec / "main")
    fpa = 120 // info["server_fps"]
    sfpa = info["shadow_fps"] // info["server_fps"]
    out = tmp_path / "out"
    print(f"case: splinter mode {mode} fpa {fpa}", flush=True)
    res = subprocess.run([str(SPLINTER), str(rec), str(out), "-w", "-m", str(mode)], capture_output=True, text=True,
                         timeout=120)
    assert res.returncode == 0, res.stderr
    n_frames = 2 * fpa
    assert sorted(p.name for p in out.glob("*.png")) == [f"{i:04d}.png" for i in range(n_frames)]

    r = R.Renderer(0)
    buf = torch.zeros((cfg.height, cfg.width, 4), dtype=torch.uint8, device="cuda")
    for pair in range(2):
        fwd, bwd = RC.read_view(rec / "main", pair, 0), RC.read_view(rec / "main", pair, 1)
        fd = R.FrameData(R.KeyframeView.from_numpy(fwd), R.KeyframeView.from_numpy(bwd))
        for b in range(fpa):
            t = b / fpa
            _, vp, ivp = RC.camera_at(fwd["view"], bwd["view"], fwd["proj"], t)
            ri = R.RenderInfo((cfg.width, cfg.height), t, vp, ivp,
                              R.AdditionalRenderInfo(fpa, sfpa, False, R.RenderMode(mode)))
            r.render(buf, ri, fd)
            torch.cuda.synchronize()
            got = RC.decode_png(out / f"{pair * fpa + b:04d}.png")
            exp = buf.cpu().numpy()
            assert np.array_equal(got, exp), (pair, b, int((got != exp).any(-1).sum()))
    r.close()
