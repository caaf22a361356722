"""CPU-only checks of the product boundary: libdevil_hip.so loads, exports
every function include/devil_hip.h declares, and reports errors as codes
when no GPU is present; host-side argument handling mirrors R/main.R."""
import ctypes as C
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def lib():
    from devil_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        from devil_amd.build import build
        build()
    return _lib.lib()


def header_functions():
    txt = open(os.path.join(ROOT, "include", "devil_hip.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|void|const char \*)\s*(devil_\w+)\(", txt,
                                 re.M)))


def test_exports_every_declared_symbol(lib):
    names = header_functions()
    assert len(names) >= 13
    from devil_amd import _lib
    assert sorted(_lib.EXPORTED) == names
    for name in names:
        assert hasattr(lib, name), name


def test_no_gpu_error_codes(lib):
    import torch
    if torch.cuda.is_available():
        pytest.skip("GPU present")
    h = C.c_void_p()
    assert lib.devil_create(C.byref(h), 0) == 6
    assert lib.devil_strerror(6) == b"no HIP device"
    assert lib.devil_version().startswith(b"devil_hip")
    import devil_amd
    with pytest.raises(devil_amd.DevilError) as ei:
        devil_amd.fit_devil(np.ones((2, 4), np.int32), np.ones((4, 1)))
    assert ei.value.code == 6


def test_cluster_blocks_mirror_R():
    import devil_amd
    # factor(levels = unique) codes, sorted check, cumsum(rle(.)$lengths)
    ends = devil_amd.cluster_blocks(["b", "b", "a", "a", "a", "c"], 6)
    np.testing.assert_array_equal(ends, [2, 5, 6])
    with pytest.raises(ValueError, match="grouped in blocks"):
        devil_amd.cluster_blocks(["a", "b", "a"], 3)
    np.testing.assert_array_equal(devil_amd.cluster_blocks([7] * 5, 5), [5])


def test_offset_vector_mirror_R():
    import devil_amd
    v = devil_amd.compute_offset_vector(0.5, 3, [1.0, 2.0, 4.0])
    np.testing.assert_allclose(v, 0.5 + np.log([1, 2, 4]), rtol=0, atol=0)
    np.testing.assert_array_equal(devil_amd.compute_offset_vector(0, 2, None), [0, 0])


def test_api_argument_errors():
    import devil_amd
    with pytest.raises(NotImplementedError):
        devil_amd.fit_devil(np.ones((2, 4)), np.ones((4, 1)), CUDA=False)


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    """Without the built library every entry into the package raises: there
    is no CPU path to fall back to."""
    from devil_amd import _lib
    monkeypatch.setattr(_lib, "_lib", None)
    monkeypatch.setattr(_lib, "LIB_PATH", str(tmp_path / "libdevil_hip.so"))
    with pytest.raises(RuntimeError, match="no CPU fallback"):
        _lib.lib()
    import devil_amd
    with pytest.raises(RuntimeError, match="no CPU fallback"):
        devil_amd.fit_devil(np.ones((4, 8), np.int32), np.ones((8, 1)))
