"""Host-only checks of the C-ABI: the library loads without a GPU, exports every
symbol include/gpx.h declares, and rejects invalid arguments with GPX_EINVAL
before touching the device."""
import ctypes
import os
import re

import pytest

from geepee_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    text = open(os.path.join(ROOT, 'include', 'gpx.h')).read()
    return sorted(set(re.findall(r'^\s*(?:long|int)\s+(gpx_\w+)\s*\(', text, flags=re.M)))


def test_header_symbols_exported_and_bound():
    L = _lib.load()
    names = _declared()
    assert len(names) >= 20
    for n in names:
        assert hasattr(L, n), n
        assert n in _lib.SIGNATURES, n


def test_invalid_arguments_return_einval():
    L = _lib.load()
    assert L.gpx_gemm_batched(0, 0, -1, 4, 4, 1.0, None, 4, 0, None, 4, 0, 0.0, None, 4, 0, 1, 1,
                              None, 0, None) == -1
    assert L.gpx_spd_inv_batched(300, 1, None, 300, 0, None, 300, 0, None, None, None) == -1
    assert L.gpx_psi_fwd(4, 4, 99, None, None, None, None, None, None, None, None) == -1
    assert L.gpx_gauss_tilted(4, 1, None, None, None, None, 0.5, 1.0, None, None, None, None, None) == -1
    assert L.gpx_emission_tilted(4, 99, 2, None, None, None, 99, None, 2, None, 2, 0.5, 1.0, None,
                                 None, None, None, None, None, None) == -1
    # empty problems are a no-op success
    assert L.gpx_gemm_batched(0, 0, 0, 4, 4, 1.0, None, 4, 0, None, 4, 0, 0.0, None, 4, 0, 1, 1,
                              None, 0, None) == 0


def test_workspace_queries():
    L = _lib.load()
    assert L.gpx_gemm_workspace(10, 20, 3, 1) == 0
    assert L.gpx_gemm_workspace(10, 20, 3, 4) == 4 * 3 * 10 * 20
    assert L.gpx_psi_bwd_workspace(100, 10, 2) > 100 * 10 * 10
    assert L.gpx_spd_max_m() == 256
    assert L.gpx_psi_qmax() == 16


def test_library_has_no_undefined_gpx_symbols():
    """Every gpx_* symbol the library references is defined in it (a shared library
    links with unresolved symbols, which would only fail at the first call)."""
    import subprocess
    from geepee_amd import _lib
    path = _lib.lib_path() if hasattr(_lib, 'lib_path') else None
    if path is None:
        import os
        path = os.path.join(os.path.dirname(_lib.__file__), 'libgpx.so')
    out = subprocess.run(['nm', '-D', '--undefined-only', path], capture_output=True, text=True, check=True)
    undefined = [l.split()[-1] for l in out.stdout.splitlines() if 'gpx_' in l]
    assert not undefined, undefined


def test_noalloc_rejects_oversize_before_any_launch():
    """The allocation-free fused forward checks its workspace against the fixed scratch on
    the host: an oversized shape returns GPX_ETOOBIG (-3) without a device call."""
    L = _lib.load()
    cap = L.gpx_fused_fwd_noalloc_max()
    assert cap == 4 << 20
    N, M, Q, D, splits = 40000, 128, 8, 32, 16
    assert L.gpx_fused_fwd_workspace(N, M, D, splits) > cap
    assert L.gpx_fused_mm_fwd_noalloc(N, M, Q, D, 1, 1, 1, 1, 1, 1, 1, splits, 1, 1, 1, 1, None) == -3
    # invalid arguments are still GPX_EINVAL
    assert L.gpx_fused_mm_fwd_noalloc(N, M, 99, D, 1, 1, 1, 1, 1, 1, 1, splits, 1, 1, 1, 1, None) == -1
