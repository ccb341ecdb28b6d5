"""CPU checks of the C-ABI library: it loads and exports every declared symbol."""

import re
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]


def declared_symbols():
    names = set()
    for header in (ROOT / "include").glob("*.h"):
        text = header.read_text()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names.update(re.findall(r"\b(pty_[a-z0-9_]+)\s*\(", text))
    return sorted(names)


def test_header_declares_entry_points():
    names = declared_symbols()
    assert "pty_forward" in names and "pty_last_error" in names


def test_library_exports_every_declared_symbol():
    from ptyrodactyl_amd import _native
    if not _native.LIB_PATH.exists():
        pytest.skip("HIP library not built")
    lib = _native.load()
    missing = [n for n in declared_symbols() if not hasattr(lib, n)]
    assert not missing, f"missing exports: {missing}"
    assert set(declared_symbols()) <= set(_native.SIGNATURES), "ctypes signatures lag the header"
    assert lib.pty_abi_version() == _native.ABI_VERSION


def test_invalid_arguments_rejected_without_gpu():
    """Argument validation runs on the host before any device call."""
    from ptyrodactyl_amd import _native
    if not _native.LIB_PATH.exists():
        pytest.skip("HIP library not built")
    lib = _native.load()
    rc = lib.pty_forward(100, 1, 1, 1, None, 100, 100, None, None, None, None, None, None, None, None, None, 0, None)
    assert rc == _native.PTY_ERR_INVALID
    assert b"window size" in lib.pty_last_error()
    with pytest.raises(ValueError, match="window size"):
        _native.check(rc, "pty_forward")
    assert lib.pty_forward_workspace_bytes(256, 10, 1, 4) == 4 * 256 * 256 * 8
    assert lib.pty_forward_workspace_bytes(128, 10, 1, 0) == 0


def test_annular_and_adjoint_queries_validate_on_host():
    from ptyrodactyl_amd import _native
    if not _native.LIB_PATH.exists():
        pytest.skip("HIP library not built")
    lib = _native.load()
    assert lib.pty_annular_integrate(None, 4, 0, None, None, None) == _native.PTY_ERR_INVALID
    assert lib.pty_annular_integrate(None, 0, 16, None, None, None) == _native.PTY_OK
    assert lib.pty_annular_integrate(None, 2, 16, None, None, None) == _native.PTY_ERR_INVALID
    assert b"null pointer" in lib.pty_last_error()
    # the gathered adjoint needs room for the contributions of every slice
    assert lib.pty_adjoint_workspace_bytes(128, 16, 4, 1024, 1) > lib.pty_adjoint_workspace_bytes(128, 4, 4, 1024, 1)
    assert lib.pty_adjoint_workspace_bytes(100, 16, 4, 1024, 1) == 0
    assert lib.pty_persistent_workspace_bytes(128, 2) == 0


def test_potential_builders_validate_on_host():
    from ptyrodactyl_amd import _native
    if not _native.LIB_PATH.exists():
        pytest.skip("HIP library not built")
    lib = _native.load()
    assert lib.pty_projected_potential(5, None, 0, None, None, None) == _native.PTY_ERR_INVALID
    assert b"parameterization" in lib.pty_last_error()
    assert lib.pty_projected_potential(0, None, 4, None, None, None) == _native.PTY_ERR_INVALID
    assert lib.pty_atom_potential_grid(0, 1, None, 0.1, 0, 8, 4, None, None, None) == _native.PTY_ERR_INVALID
    assert b"height and width" in lib.pty_last_error()
    assert lib.pty_atom_potential_grid(1, 1, None, 0.1, 8, 8, 0, None, None, None) == _native.PTY_ERR_INVALID
    assert b"supersampling" in lib.pty_last_error()
    assert lib.pty_atom_potential_grid(1, 1, None, -0.1, 8, 8, 2, None, None, None) == _native.PTY_ERR_INVALID
    assert b"pixel_size" in lib.pty_last_error()
    assert lib.pty_potential_slices_fourier(1, 1, 8, 8, -1, None, None, None, None, None) == _native.PTY_ERR_INVALID
    assert b"num_atoms" in lib.pty_last_error()
    assert lib.pty_potential_slices_fourier(2, 1, 8, 8, 3, None, None, None, None, None) == _native.PTY_ERR_INVALID
    assert b"null pointer" in lib.pty_last_error()
    assert lib.pty_atom_center_grad_workspace_bytes(3) == 3 * 512 * 16
    assert lib.pty_atom_potential_grid_center_grad(5, 1, None, 0.1, 8, 8, 1, None, None, None, None, 0,
                                                   None) == _native.PTY_ERR_INVALID
    assert b"parameterization" in lib.pty_last_error()
    assert lib.pty_atom_potential_grid_center_grad(0, 1, None, 0.1, 8, 8, 1, None, None, None, None, 0,
                                                   None) == _native.PTY_ERR_INVALID
    assert b"null pointer" in lib.pty_last_error()


def test_grid_entry_points_validate_on_host():
    """Runtime-size entry points: edge limits, seed exclusivity and workspace sizes, checked before any launch."""
    from ptyrodactyl_amd import _native
    if not _native.LIB_PATH.exists():
        pytest.skip("HIP library not built")
    lib = _native.load()
    args = (1, 1, 1, None, 100, 100, None, None, None, None, None, None)
    rc = lib.pty_forward_grid(5000, 100, *args, None, None, None, 1, None, 0, None)
    assert rc == _native.PTY_ERR_INVALID and b"each edge must be in [1, 4096]" in lib.pty_last_error()
    rc = lib.pty_forward_grid(0, 100, *args, None, None, None, 1, None, 0, None)
    assert rc == _native.PTY_ERR_INVALID
    rc = lib.pty_forward_grid(100, 100, *args, None, None, None, 1, None, 0, None)
    assert rc == _native.PTY_ERR_INVALID and b"null input pointer" in lib.pty_last_error()
    # one forward workgroup holds one h x w complex64 wave (plus the streaming passes' twiddle tables,
    # (h + w) x 8 bytes rounded up to 256); the adjoint k + 2 waves (k clamped to S), one more wave for the
    # streaming adjoint's slab-gradient contributions, 16 bytes of shift partials per row group (one group of
    # 15 rows here) and 256 of alignment
    tw = 512
    assert lib.pty_grid_workspace_bytes(15, 32, 8, 4, 1, 0) == 15 * 32 * 8 + tw
    assert lib.pty_grid_workspace_bytes(15, 32, 8, 4, 3, 1) == 3 * (15 * 32 * 8 * 7 + 16) + tw + 256
    assert lib.pty_grid_workspace_bytes(15, 32, 2, 4, 1, 1) == 15 * 32 * 8 * 5 + 16 + tw + 256
    assert lib.pty_grid_workspace_bytes(15, 32, 8, 4, 10 ** 6, 0) == 2048 * 15 * 32 * 8 + tw
    assert lib.pty_grid_workspace_bytes(4097, 32, 8, 4, 1, 0) == 0
    fake = 16  # any non-null address: validation fails before it is used
    rc = lib.pty_adjoint_grid(15, 32, 1, 1, 1, fake, 15, 32, fake, fake, fake, fake, fake, fake, fake, fake, None, 1,
                              fake, None, None, fake, 1 << 20, None)
    assert rc == _native.PTY_ERR_INVALID and b"exactly one of grad_patterns and grad_amplitude" in lib.pty_last_error()
    rc = lib.pty_adjoint_grid(15, 32, 1, 1, 1, fake, 15, 32, fake, fake, fake, fake, fake, fake, fake, None, None, 1,
                              fake, None, None, fake, 100, None)
    assert rc == _native.PTY_ERR_INVALID and b"workspace" in lib.pty_last_error()


def test_grid_workspace_holds_every_mode_of_a_position():
    """The streaming runtime-size kernels (grids of 2^14 pixels or more) keep the M waves of a position in
    flight: the engine's workspace for one position and two modes must hold them (plus the twiddle tables), or
    the launcher falls back to one workgroup per position."""
    from ptyrodactyl_amd import _native, engine
    if not _native.LIB_PATH.exists():
        pytest.skip("HIP library not built")
    h, w, s_count, k = 150, 130, 5, 2
    wave = h * w * 8
    twiddles = ((h + w) * 8 + 255) // 256 * 256
    fwd = engine.grid_workspace_bytes(h, w, s_count, 1, 1, 2, adjoint=False)
    assert fwd >= twiddles + 2 * wave
    adj = engine.grid_workspace_bytes(h, w, s_count, k, 1, 2, adjoint=True)
    assert adj >= twiddles + 256 + 2 * (k + 2) * wave


def test_sqrt_loss_validates_on_host():
    from ptyrodactyl_amd import _native
    if not _native.LIB_PATH.exists():
        pytest.skip("HIP library not built")
    lib = _native.load()
    assert lib.pty_sqrt_intensity_loss_workspace_bytes() == 1024 * 8
    assert lib.pty_sqrt_intensity_loss(4, None, None, 1e-6, None, None, None, 0, None) == _native.PTY_ERR_INVALID
    assert b"null pointer" in lib.pty_last_error()
    fake = 1 << 20  # never dereferenced: the workspace check fails first
    rc = lib.pty_sqrt_intensity_loss(4, fake, fake, 1e-6, None, None, fake, 8, None)
    assert rc == _native.PTY_ERR_INVALID and b"workspace" in lib.pty_last_error()
    assert lib.pty_sqrt_intensity_loss(-1, fake, fake, 1e-6, None, None, fake, 8192, None) == _native.PTY_ERR_INVALID
