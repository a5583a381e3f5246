"""The C-ABI library loads without a GPU and exports every declared entry point."""
import ctypes
import os
import re

import numpy as np

from wavefrontpathtracing_amd import _lib as L

import oracle_api as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    return sorted(set(re.findall(r"WF_API\s+[\w\s\*]+?\b(wfs?_\w+)\s*\(", text)))


def test_every_declared_symbol_is_exported():
    lib = ctypes.CDLL(L.LIB_PATH)
    names = _declared("wf.h") + _declared("wf_scene.h")
    assert len(names) > 30
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_table_matches_header():
    names = set(_declared("wf.h") + _declared("wf_scene.h"))
    assert names == set(L.SIGNATURES), (names ^ set(L.SIGNATURES))


def test_host_only_entry_points_run_without_gpu():
    L.lib()   # loads
    import numpy as np
    from wavefrontpathtracing_amd import renderer as R
    blob = R.bvh_build(np.arange(9, dtype=np.float32).reshape(1, 9) * np.float32(0.1))
    assert len(blob) == 12 + 2 * 32 + 36 + 4


def test_oracle_loads():
    assert O.olib() is not None


def test_image_files_round_trip(tmp_path):
    from wavefrontpathtracing_amd import renderer as R
    rng = np.random.default_rng(2)
    img = rng.random((5, 7, 4)).astype(np.float32)
    R.write_pfm(tmp_path / "a.pfm", img)
    raw = (tmp_path / "a.pfm").read_bytes()
    head = b"PF\n7 5\n-1.0\n"
    assert raw.startswith(head)
    data = np.frombuffer(raw[len(head):], dtype="<f4").reshape(5, 7, 3)
    assert np.array_equal(data[::-1], img[..., :3])          # bottom row first
    img8 = rng.integers(0, 256, (3, 4, 4), dtype=np.uint8)
    R.write_ppm(tmp_path / "b.ppm", img8)
    raw = (tmp_path / "b.ppm").read_bytes()
    head = b"P6\n4 3\n255\n"
    assert raw.startswith(head)
    assert np.array_equal(np.frombuffer(raw[len(head):], dtype=np.uint8).reshape(3, 4, 3), img8[..., :3])
