"""The splinter driver's frame schedule, checked on the CPU with --dry-run
(no device is touched): which keyframe pair each client frame renders, its
alpha, when the renderer is handed a new pair, and the target camera. The
expected schedule restates the reference's loop (src/main.cu:149-191): read
the next pair when latency == b, alpha = (latency > b ? fpa + b : b) / fpa,
read again after the last frame of a round when latency > b; the camera is
glm::interpolate of the pair's views, or frames[n] of a -c camera file with
its clientFps setting frames_per_anchor (main.cu:128-161)."""
import subprocess
from pathlib import Path

import numpy as np
import pytest

from imagebasedspatiotemporalinterpolation_amd import recording as RC
from imagebasedspatiotemporalinterpolation_amd import scene as S

SPLINTER = Path(__file__).resolve().parent.parent / "imagebasedspatiotemporalinterpolation_amd" / "splinter"
N_KEYFRAMES = 5


def reference_schedule(total, fpa, latency, hold=False):
    """[(frame, pair or None, alpha, rendered)] following src/main.cu:149-191.
    Frames before the first pair is read are not rendered."""
    reads, out = 0, []
    for i in range(0, total, fpa):
        for b in range(fpa):
            if latency == b:
                reads += 1
            t = np.float32(fpa + b if latency > b else b) / np.float32(fpa)
            n = i + b
            if n + 1 > total:
                break
            rendered = reads > 0 and not (b < latency and hold)
            out.append((n, reads - 1 if reads else None, float(t), rendered))
            if latency > b and b == fpa - 1:
                reads += 1
    return out


def dry_run(*args):
    res = subprocess.run([str(SPLINTER), *map(str, args), "--dry-run"], capture_output=True, text=True, timeout=120)
    assert res.returncode == 0, res.stderr
    rows = []
    for line in res.stdout.splitlines():
        f = line.split()
        if not f or f[0] != "frame":
            continue
        if f[4] == "skip":
            rows.append((int(f[1]), None, None, None, None))
        else:
            vp = np.array([float(x) for x in f[11:27]], np.float32)
            rows.append((int(f[1]), int(f[5]), float(f[7]), int(f[9]), vp))
    return rows


@pytest.fixture(scope="module")
def recording(tmp_path_factory):
    cfg = S.defaults(1, 5)
    cfg.width, cfg.height = 64, 36
    root = tmp_path_factory.mktemp("rec")
    RC.write_recording(root / "rec", cfg, N_KEYFRAMES, shadows=False)
    info, frames = RC.read_camera(root / "rec" / "main")
    return cfg, root, info


def check(rows, expect, camera):
    assert [r[0] for r in rows] == [e[0] for e in expect]
    last_pair = None
    for (n, pair, t, new, vp), (en, epair, et, rendered) in zip(rows, expect):
        if not rendered:
            assert pair is None, n
            continue
        assert (pair, t) == (epair, et), (n, pair, t, epair, et)
        # the renderer is handed a pair exactly at the first frame rendered from it
        assert new == int(pair != last_pair), (n, new, pair, last_pair)
        last_pair = pair
        np.testing.assert_array_equal(vp, camera(n, pair, t))


@pytest.mark.parametrize("latency,hold", [(0, False), (1, False), (3, True), ("fpa", False), ("fpa", True)])
def test_recording_schedule_with_latency(recording, tmp_path, latency, hold):
    cfg, root, info = recording
    fpa = 120 // info["server_fps"]
    lat = fpa if latency == "fpa" else latency
    total = 3 * fpa
    args = [root / "rec", tmp_path / "out", "-l", lat, "--frames", total] + (["-lh"] if hold else [])
    rows = dry_run(*args)
    views = {}

    def camera(n, pair, t):
        if pair not in views:
            views[pair] = (RC.read_view(root / "rec" / "main", pair, 0), RC.read_view(root / "rec" / "main", pair, 1))
        f, b = views[pair]
        return RC.camera_at(f["view"], b["view"], f["proj"], t)[1]

    expect = reference_schedule(total, fpa, lat, hold)
    if hold and lat >= fpa:  # every frame waits for a late keyframe: nothing is rendered
        assert not any(e[3] for e in expect)
    else:
        assert sum(e[3] for e in expect) > fpa  # frames from at least two pairs
    check(rows, expect, camera)


def test_recording_schedule_with_client_camera_file(recording, tmp_path):
    """-c: the file's frame count is the run length and its clientFps (60 here,
    the recording's server runs at 15 Hz) gives frames_per_anchor 4; frame n
    renders with normalizeProjMat(frames[n].proj) * frames[n].w2c."""
    cfg, root, info = recording
    client_fps = 60
    fpa = client_fps // info["server_fps"]
    n_frames = 3 * fpa + 2
    cam_dir = RC.write_camera_path(tmp_path / "client", cfg, n_frames, client_fps, frame_time=0.75)
    _, frames = RC.read_camera(cam_dir)
    rows = dry_run(root / "rec", tmp_path / "out", "-c", cam_dir, "-l", 1)
    expect = reference_schedule(n_frames, fpa, 1)
    check(rows, expect, lambda n, pair, t: RC.client_camera(frames[n][1], frames[n][2])[0])


def test_synthetic_scene_with_client_camera_file(tmp_path):
    cfg = S.defaults(1, 5)
    cam_dir = RC.write_camera_path(tmp_path / "client", cfg, 7, 120, frame_time=0.5)
    _, frames = RC.read_camera(cam_dir)
    rows = dry_run("scene:1:5", tmp_path / "out", "-c", cam_dir)
    fpa = cfg.keyframe_spacing
    check(rows, reference_schedule(7, fpa, 0), lambda n, pair, t: RC.client_camera(frames[n][1], frames[n][2])[0])


def test_client_camera_is_normalised_projection_times_view():
    """viewproj = U * D * R * proj * w2c with R the reversed-Z flip, D the
    [0,1] -> [-1,1] depth map and U Unity's y flip (file_reader.cu:219-230,
    file_reader.h:60-67), checked in float64."""
    rng = np.random.default_rng(2)
    w2c = rng.standard_normal((4, 4))
    w2c[3] = [0, 0, 0, 1]
    proj = np.zeros((4, 4))
    proj[0, 0], proj[1, 1], proj[2, 2], proj[2, 3], proj[3, 2] = 1.3, 1.7, 0.01, 0.3, -1.0
    R = np.array([[1, 0, 0, 0], [0, 1, 0, 0], [0, 0, -1, 1], [0, 0, 0, 1]], float)
    D = np.array([[1, 0, 0, 0], [0, 1, 0, 0], [0, 0, 2, -1], [0, 0, 0, 1]], float)
    U = np.diag([1.0, -1.0, 1.0, 1.0])
    expect = U @ D @ R @ proj @ w2c
    cm = lambda m: np.asarray(m, np.float32).T.reshape(16)
    vp, ivp = RC.client_camera(cm(w2c), cm(proj))
    got = vp.reshape(4, 4).T.astype(np.float64)
    np.testing.assert_allclose(got, expect, rtol=1e-5, atol=1e-5)
    np.testing.assert_allclose(got @ ivp.reshape(4, 4).T.astype(np.float64), np.eye(4), atol=1e-4)


def test_camera_path_round_trip(tmp_path):
    cfg = S.defaults(1, 3)
    d = RC.write_camera_path(tmp_path / "c", cfg, 5, 90, frame_time=2.0)
    info, frames = RC.read_camera(d)
    assert (info["client_fps"], info["server_fps"], info["n_frames"]) == (90, 10, 5)
    for n, (idx, w2c, proj) in enumerate(frames):
        vm, pm = S.camera_parts(cfg, 2.0 * n)
        assert idx == n
        np.testing.assert_array_equal(w2c, vm)
        np.testing.assert_array_equal(proj, RC.raw_projection(pm))
