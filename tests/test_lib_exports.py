"""The C-ABI library builds for gfx950, loads, and exports every symbol include/vamx.h declares (CPU)."""

import os
import re
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    src = open(os.path.join(ROOT, "include", "vamx.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:int|void|const char\*)\s+(vamx_\w+)\s*\(", src, flags=re.M)))


def test_header_declares_entry_points():
    names = _declared()
    assert "vamx_par_forward" in names and "vamx_rt_deposit" in names and "vamx_bclp_step" in names
    assert len(names) >= 17


def test_library_exports_every_declared_symbol():
    from vamtoolbox_amd import _lib
    lib_path = _lib.LIB_PATH
    assert os.path.exists(lib_path), "run python -m vamtoolbox_amd.build first"
    out = subprocess.run(["nm", "-D", "--defined-only", lib_path], capture_output=True, text=True, check=True).stdout
    exported = set(re.findall(r"\s[TW]\s(vamx_\w+)$", out, flags=re.M))
    missing = [n for n in _declared() if n not in exported]
    assert not missing, missing
    # the ctypes binding knows every declared symbol and the library satisfies it
    assert set(_lib.SIGNATURES) == set(_declared())
    _lib.load()
    assert _lib.MISSING == []
    assert _lib.load().vamx_version() == 1


def test_library_targets_gfx950():
    from vamtoolbox_amd import _lib
    out = subprocess.run(["strings", _lib.LIB_PATH], capture_output=True, text=True).stdout
    assert "gfx950" in out


def test_error_path_without_gpu_work():
    """Argument validation returns a status and a message (no device work needed)."""
    import ctypes
    from vamtoolbox_amd import _lib
    lib = _lib.load()
    h = ctypes.c_void_p()
    rc = lib.vamx_par_create(None, ctypes.byref(h))
    assert rc == -1
    assert b"null" in lib.vamx_last_error()
    rc = lib.vamx_par_forward(None, None, None, None)
    assert rc == -3


def test_order_stats_argument_errors_without_gpu_work():
    """vamx_order_stats validates its arguments before any launch."""
    import ctypes
    from vamtoolbox_amd import _lib
    lib = _lib.load()
    nb = ctypes.c_int64()
    assert lib.vamx_order_stats_workspace_bytes(ctypes.byref(nb)) == 0 and nb.value >= 8 * 256 * 8
    assert lib.vamx_order_stats_workspace_bytes(None) == -1
    fake = ctypes.c_void_p(256)   # never dereferenced: every call below fails validation first
    assert lib.vamx_order_stats(None, 10, 0, fake, fake, None) == -1
    assert lib.vamx_order_stats(fake, 0, 0, fake, fake, None) == -1 and b"empty" in lib.vamx_last_error()
    assert lib.vamx_order_stats(fake, 10, 10, fake, fake, None) == -1 and b"rank" in lib.vamx_last_error()
    assert lib.vamx_order_stats(fake, 10, -1, fake, fake, None) == -1
    assert lib.vamx_order_stats(fake, 10, 3, fake, ctypes.c_void_p(260), None) == -1
    assert b"aligned" in lib.vamx_last_error()


def test_osmo_elementwise_argument_errors_without_gpu_work():
    """vamx_osmo_model_update / vamx_scale_clamp validate their arguments and
    launch nothing for an empty array."""
    import ctypes
    from vamtoolbox_amd import _lib
    lib = _lib.load()
    fake = ctypes.c_void_p(256)   # never dereferenced
    assert lib.vamx_osmo_model_update(None, fake, fake, 10, 0.5, 0, None) == -1
    assert lib.vamx_osmo_model_update(fake, fake, fake, -1, 0.5, 0, None) == -1
    assert lib.vamx_osmo_model_update(fake, fake, fake, 0, 0.5, 1, None) == 0
    assert lib.vamx_scale_clamp(fake, 10, None, 0.0, None) == -1
    assert lib.vamx_scale_clamp(fake, -3, fake, 0.0, None) == -1
    assert lib.vamx_scale_clamp(fake, 0, fake, 0.0, None) == 0
    # vamx_scale_clamp_segments(v, n, seg_off, k, lo, workspace, stream)
    assert lib.vamx_scale_clamp_segments(fake, 10, None, 2, 0.0, fake, None) == -1
    assert lib.vamx_scale_clamp_segments(fake, 10, fake, 0, 0.0, fake, None) == -1
    assert lib.vamx_scale_clamp_segments(fake, 10, fake, 70000, 0.0, fake, None) == -1
    assert lib.vamx_scale_clamp_segments(fake, -1, fake, 2, 0.0, fake, None) == -1
    assert lib.vamx_scale_clamp_segments(fake, 10, ctypes.c_void_p(260), 2, 0.0, fake, None) == -1
    assert lib.vamx_scale_clamp_segments(fake, 10, fake, 2, 0.0, ctypes.c_void_p(258), None) == -1
    assert lib.vamx_scale_clamp_segments(fake, 0, fake, 2, 0.0, fake, None) == 0


def test_ver_argument_errors_without_gpu_work():
    """vamx_ver validates its arguments before any launch."""
    import ctypes
    from vamtoolbox_amd import _lib
    lib = _lib.load()
    nb = ctypes.c_int64()
    assert lib.vamx_ver_workspace_bytes(ctypes.byref(nb)) == 0 and nb.value >= 16
    fake = ctypes.c_void_p(256)   # never dereferenced
    assert lib.vamx_ver(None, fake, fake, 10, 5, fake, fake, None) == -1
    assert lib.vamx_ver(fake, fake, fake, 0, 5, fake, fake, None) == -1
    assert lib.vamx_ver(fake, fake, fake, 10, 0, fake, fake, None) == -1
    assert lib.vamx_ver(fake, fake, fake, 10, 5, fake, ctypes.c_void_p(260), None) == -1
