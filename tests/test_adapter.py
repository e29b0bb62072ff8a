"""The reference-side adapter (qoraholo_amd/adapter/hip_backend.cpp: Backend::hip behind the
reference's dispatch, with the signatures of its cuda:: backend, src/backends.hpp:107-126)
compiles against the reference's own headers, and every C-ABI function it calls is exported by
libqoraholo_hip.so. Skipped where the reference's sources are absent."""

import os
import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
REF = Path(os.environ.get("QH_REFERENCE_ROOT", "/root/reference"))
ADAPTER = ROOT / "qoraholo_amd" / "adapter" / "hip_backend.cpp"

needs_reference = pytest.mark.skipif(
    not (REF / "src" / "backends.hpp").exists() or shutil.which("g++") is None,
    reason="the reference's headers or g++ are not available")


def _compile(tmp_path: Path) -> Path:
    obj = tmp_path / "hip_backend.o"
    cmd = ["g++", "-std=c++20", "-Wall", "-Wextra", "-Werror", "-fPIC", "-c",
           "-I", str(REF / "include"), "-I", str(REF / "src"), "-I", str(ROOT / "include"),
           str(ADAPTER), "-o", str(obj)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    assert res.returncode == 0, res.stderr
    return obj


def _symbols(args) -> set[str]:
    out = subprocess.run(["nm", *args], capture_output=True, text=True, check=True).stdout
    return {line.split()[-1] for line in out.splitlines() if line.strip()}


@needs_reference
def test_adapter_compiles_against_reference_headers(tmp_path):
    obj = _compile(tmp_path)
    assert obj.stat().st_size > 0


@needs_reference
def test_adapter_calls_only_exported_entry_points(tmp_path):
    from qoraholo_amd import _native

    obj = _compile(tmp_path)
    called = {s for s in _symbols(["-u", str(obj)]) if s.startswith("qh_")}
    exported = _symbols(["-D", "--defined-only", str(_native.LIB_PATH)])
    assert len(called) >= 15
    assert called <= exported, sorted(called - exported)
    # Every table-taking BPM call of the reference's backend ABI is among them.
    for name in ("qh_hip_bpm_forward_tables", "qh_hip_bpm_vjp_tables",
                 "qh_hip_bpm_reconstruct_tables", "qh_hip_bpm_jvp_tables",
                 "qh_hip_pipeline_create", "qh_hip_reconstruct"):
        assert name in called, name
