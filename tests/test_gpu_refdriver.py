"""The reference's own 1D shock-tube program (cans1d/md_mhdshktb/hdmlw:
main.f, model.f and bnd.f, unchanged, compiled from the reference sources by
oracle/Makefile `drivers`) linked once with the reference routines and once
with the drop-in library libcans1d_gpu.so in place of libcans1d.a.  Its
mlw_m, cfl_m, bdsppx and bdsmnx calls then run on the GPU through the
strict host-array policy (every call copies its arrays in and out), and every
file the program writes (the DAC records of t, ro, pr, vx, vy, by at each
output time, x, bx, params.txt, out.txt) must be byte-identical."""
import os
import subprocess
import time
from pathlib import Path

import pytest

pytestmark = pytest.mark.gpu

ROOT = Path(__file__).resolve().parents[1]
REF = ROOT / "oracle" / "_ref" / "shktb1d_ref"
GPU = ROOT / "oracle" / "_ref" / "shktb1d_gpu"


def _run(exe: Path, cwd: Path) -> float:
    cwd.mkdir(parents=True, exist_ok=True)
    t0 = time.perf_counter()
    p = subprocess.run([str(exe)], cwd=cwd, capture_output=True, text=True, timeout=100)
    sec = time.perf_counter() - t0
    assert p.returncode == 0, (exe.name, p.returncode, p.stdout[-2000:], p.stderr[-2000:])
    assert "normal stop" in p.stdout, p.stdout[-2000:]
    return sec


def test_reference_main_on_dropin_library(dev, tmp_path):
    if not (REF.exists() and GPU.exists()):
        pytest.skip("reference driver programs not built (oracle/Makefile drivers needs the reference sources)")
    print("CASE reference md_mhdshktb/hdmlw main.f: reference routines vs libcans1d_gpu.so", flush=True)
    t_ref = _run(REF, tmp_path / "ref")
    t_gpu = _run(GPU, tmp_path / "gpu")
    files = sorted(os.listdir(tmp_path / "ref"))
    assert "ro.dac" in files and "t.dac" in files, files
    assert sorted(os.listdir(tmp_path / "gpu")) == files
    for name in files:
        a = (tmp_path / "ref" / name).read_bytes()
        b = (tmp_path / "gpu" / name).read_bytes()
        assert a == b, f"{name} differs ({len(a)} vs {len(b)} bytes)"
    steps = int((tmp_path / "ref" / "out.txt").read_text().split("stop")[1].split("step=")[1].split()[0])
    print(f"STRICT drop-in: {steps} steps of 1024 cells; reference build {t_ref:.3f} s, "
          f"GPU drop-in (strict, PCIe per call) {t_gpu:.3f} s = {t_gpu / steps * 1e6:.1f} us/step", flush=True)
