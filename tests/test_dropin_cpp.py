"""The drop-in from C++: code written against the reference API, compiled
against include/goofy_tc.h and linked with libgoofy_hip.so (built by
goofy_amd/build.py into tests/native/libdropin_example.so), produces the
reference's blocks."""
import ctypes
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
SO = ROOT / "tests" / "native" / "libdropin_example.so"


def load():
    if not SO.exists():
        pytest.fail(f"{SO} not built (python goofy_amd/build.py)")
    lib = ctypes.CDLL(str(SO))
    lib.dropin_main.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_char_p)]
    return lib


def test_dropin_program_links_the_library():
    lib = load()
    assert hasattr(lib, "dropin_main")


@pytest.mark.gpu
def test_dropin_program_output_matches_reference(gpu, golden_img512, tmp_path):
    lib = load()
    args = (ctypes.c_char_p * 3)(b"dropin_example", str(tmp_path / "d.bin").encode(), str(tmp_path / "e.bin").encode())
    assert lib.dropin_main(3, args) == 0
    assert np.array_equal(np.fromfile(tmp_path / "d.bin", np.uint8), golden_img512["dxt1"])
    assert np.array_equal(np.fromfile(tmp_path / "e.bin", np.uint8), golden_img512["etc1"])
