"""The C-ABI libraries load and export every symbol their headers declare
(no device calls: these run without a GPU)."""
import ctypes
import re
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent


def _declared(header: Path):
    text = header.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(isti_[a-z_0-9]+)\s*\(", text)))


@pytest.mark.parametrize("header,lib", [
    ("include/isti.h", "imagebasedspatiotemporalinterpolation_amd/libisti.so"),
    ("include/isti_learned.h", "imagebasedspatiotemporalinterpolation_amd/libisti.so"),
    ("include/isti_quality.h", "imagebasedspatiotemporalinterpolation_amd/libisti.so"),
    ("include/isti_codec.h", "imagebasedspatiotemporalinterpolation_amd/libisti.so"),
    ("include/isti_io.h", "imagebasedspatiotemporalinterpolation_amd/libisti_io.so"),
    ("include/isti_scene.h", "imagebasedspatiotemporalinterpolation_amd/libisti_scene.so"),
])
def test_exports(header, lib):
    names = _declared(ROOT / header)
    assert names, header
    L = ctypes.CDLL(str(ROOT / lib))
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_python_binding_matches_header():
    from imagebasedspatiotemporalinterpolation_amd import _native as N

    assert sorted(N.EXPORTED_SYMBOLS) == _declared(ROOT / "include/isti.h")
    # struct sizes as the C compiler lays them out (x86-64 SysV)
    assert ctypes.sizeof(N.ViewDesc) == 8 + 8 * 8 + 64 + 64 + 8
    assert ctypes.sizeof(N.RenderInfoC) == 12 + 128 + 20
    assert ctypes.sizeof(N.ScratchView) == 8 + 24


def test_version_string():
    from imagebasedspatiotemporalinterpolation_amd import _native as N

    assert N.lib().isti_version().decode().startswith("isti-gfx950")


def test_create_without_device_reports_error():
    """isti_create on a machine without a HIP device returns a status, never throws."""
    import torch

    if torch.cuda.is_available():
        pytest.skip("a device is present")
    from imagebasedspatiotemporalinterpolation_amd import _native as N

    h = ctypes.c_void_p()
    cfg = N.Config(0, 0)
    st = N.lib().isti_create(ctypes.byref(cfg), ctypes.byref(h))
    assert st == N.ISTI_ERR_DEVICE
    assert N.lib().isti_last_error(None)


def test_learned_binding_matches_header():
    """Every entry point of include/isti_learned.h has ctypes argument types."""
    from imagebasedspatiotemporalinterpolation_amd import learned as Ln

    L = Ln._lib()
    for n in _declared(ROOT / "include/isti_learned.h"):
        assert getattr(L, n).argtypes is not None, n
    n = int(L.isti_unet_param_count())
    assert n == sum(o[3] + int(__import__("numpy").prod(o[1])) for o in Ln.param_layout())


def test_checkpoint_format_rejects_foreign_files(tmp_path):
    import numpy as np

    from imagebasedspatiotemporalinterpolation_amd import learned as Ln

    p = tmp_path / "x.bin"
    p.write_bytes(b"not a checkpoint")
    with pytest.raises(ValueError):
        Ln.load_params(p)
    Ln.save_params(p, np.zeros(5, np.float32), {})
    with pytest.raises(ValueError, match="parameters"):
        Ln.load_params(p)
