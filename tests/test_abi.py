"""The C-ABI library loads without a GPU and exports every entry point that
include/cans_gpu.h declares; the Python mirror validates its arguments
before any device call."""
import ctypes as C
import re
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
HEADER = ROOT / "include" / "cans_gpu.h"


def declared_functions():
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?[a-z_]+\s*\*?\s*(cans[0-9a-z_]+)\s*\(", text, flags=re.M)
    return sorted(set(names))


def test_header_declares_the_hot_path():
    names = declared_functions()
    for must in ("cans1d_mlw_m", "cans2d_mlw_m", "cans3d_mlw_m", "cans1d_roe_m", "cans2d_roe_m", "cans3d_roe_m",
                 "cans1d_cfl_m", "cans2d_cfl_m", "cans3d_cfl_m", "cans1d_bnd", "cans2d_bnd", "cans3d_bnd",
                 "cans3d_hlld_engine", "cans_run", "cans_gpu_map"):
        assert must in names, must


def test_library_exports_every_declared_symbol():
    from utrip_amd import LIB_GPU
    if not LIB_GPU.exists():
        pytest.skip("libcans_gpu.so not built")
    lib = C.CDLL(str(LIB_GPU))
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_fortran_dropin_libraries_export_reference_names():
    from utrip_amd import PKG_DIR
    helpers = ["mlwhalf_", "mlwfull_", "mlwartv_", "mlwsrch_", "mlwsrcf_", "roeflux_m_", "tvdminmod_"]
    want = {
        "1d": ["mlw_m_", "roe_m_", "cfl_m_", "bdfrex_", "bdperx_", "bdsppx_", "bdspnx_", "bdsmpx_", "bdsmnx_",
               "bdcnsx_", "bdfrdx_"] + helpers,
        "2d": ["mlw_m_", "roe_m_", "cfl_m_", "bdperx_", "bdpery_", "bdfrex_", "bdfrey_"] + helpers,
        "3d": ["mlw_m_", "roe_m_", "cfl_m_", "bdperx_", "bdpery_", "bdperz_", "engine_", "hlld_mh_",
               "hlld_flux_"] + helpers,
    }
    found = False
    for dim, names in want.items():
        p = PKG_DIR / f"libcans{dim}_gpu.so"
        if not p.exists():
            continue
        found = True
        lib = C.CDLL(str(p))
        missing = [n for n in names if not hasattr(lib, n)]
        assert not missing, (dim, missing)
    if not found:
        pytest.skip("Fortran drop-in libraries not built")


def test_tuning_switches_validate_without_a_gpu():
    """The two host-side tuning switches take effect without a device call
    and reject values outside their range with the library's error."""
    pytest.importorskip("torch")
    from utrip_amd import LIB_GPU, cans
    if not LIB_GPU.exists():
        pytest.skip("libcans_gpu.so not built")
    for bad in (-1, 2):
        with pytest.raises(cans.CansError, match="mlw3d path"):
            cans.set_mlw3d_path(bad)
    for good in (0, 1):
        cans.set_mlw3d_path(good)
    with pytest.raises(cans.CansError, match="march waves"):
        cans.set_march_waves(0)
    cans.set_march_waves(32768)


def test_python_mirror_rejects_bad_arrays():
    torch = pytest.importorskip("torch")
    from utrip_amd import cans
    a = torch.zeros(10, dtype=torch.float64)
    with pytest.raises(ValueError):
        cans._ptr(a, (10,), "ro")            # host tensor
    with pytest.raises(TypeError):
        cans._ptr(a.float(), (10,), "ro")    # wrong dtype
    with pytest.raises(TypeError):
        cans._ptr([1.0], None, "ro")


def test_fortran_dropin_libraries_export_only_cans_names():
    """The drop-in libraries keep the Fortran runtime they carry local: a
    CANS main program linked with its own runtime must not bind its I/O unit
    tables to a second copy."""
    import subprocess
    from utrip_amd import PKG_DIR
    found = False
    for dim in ("1d", "2d", "3d"):
        p = PKG_DIR / f"libcans{dim}_gpu.so"
        if not p.exists():
            continue
        found = True
        out = subprocess.run(["nm", "-D", "--defined-only", str(p)], capture_output=True, text=True, check=True).stdout
        names = [ln.split()[-1] for ln in out.splitlines() if ln.strip()]
        bad = [n for n in names if not (n.endswith("_") and n == n.lower() and not n.startswith("_"))]
        assert not bad, (dim, bad[:10])
        assert "mlw_m_" in names and "cfl_m_" in names
    if not found:
        pytest.skip("Fortran drop-in libraries not built")
