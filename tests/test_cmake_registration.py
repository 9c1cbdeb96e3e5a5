"""The plugin model through the reference's own build system (CPU only).

A scratch copy of the reference tree gets the model directory as src/mi355x and
one register_model line beside the others (CMakeLists.txt:303-313); the
reference's CMake then configures with -DMODEL=mi355x (cmake/register_models.cmake:
115-146, the model's model.cmake supplying setup()/setup_target()), compiles the
reference's unchanged driver with the model and links libclover_mi355x.so into
mi355x-cloverleaf.  The binary's own --help runs without a GPU.  Skipped where the
reference tree or cmake is absent (as on machines that only run the GPU tests).
"""
import os
import shutil
import subprocess
from pathlib import Path

import pytest

REPO = Path(__file__).resolve().parent.parent
REF = Path("/root/reference")
MODEL = REPO / "cloverleaf_amd" / "model_mi355x"

pytestmark = pytest.mark.skipif(not (REF / "CMakeLists.txt").exists() or shutil.which("cmake") is None,
                                reason="needs the reference tree and cmake")


def test_registers_builds_and_links_with_the_reference_cmake(tmp_path):
    lib = REPO / "cloverleaf_amd" / "libclover_mi355x.so"
    assert lib.exists(), "build the library first (make -C cloverleaf_amd/csrc)"
    tree = tmp_path / "ref"
    top_level_builds = lambda d, names: [n for n in names if Path(d) == REF and (n in (".git", "build") or n.startswith("cmake-build"))]
    shutil.copytree(REF, tree, ignore=top_level_builds)
    shutil.copytree(MODEL, tree / "src" / "mi355x", ignore=shutil.ignore_patterns("_dropin", "Makefile", "*.o"))
    cml = tree / "CMakeLists.txt"
    text = cml.read_text()
    anchor = "register_model(sycl-usm USE_SYCL_USM ${MODEL_SRC})"
    assert anchor in text
    cml.write_text(text.replace(anchor, anchor + "\nregister_model(mi355x USE_MI355X ${MODEL_SRC})", 1))
    build = tmp_path / "build"
    env = dict(os.environ)
    cfg = subprocess.run(["cmake", "-S", str(tree), "-B", str(build), "-DMODEL=mi355x", f"-DCLOVER_MI355X_ROOT={REPO}",
                          "-DENABLE_MPI=OFF", "-DCMAKE_CXX_COMPILER=g++", "-DCMAKE_BUILD_TYPE=Release"],
                         capture_output=True, text=True, timeout=600, env=env)
    assert cfg.returncode == 0, cfg.stdout[-3000:] + cfg.stderr[-3000:]
    jobs = str(min(8, os.cpu_count() or 2))
    b = subprocess.run(["cmake", "--build", str(build), "-j", jobs], capture_output=True, text=True, timeout=1200, env=env)
    assert b.returncode == 0, b.stdout[-3000:] + b.stderr[-3000:]
    exe = build / "mi355x-cloverleaf"
    assert exe.exists()
    ldd = subprocess.run(["ldd", str(exe)], capture_output=True, text=True).stdout
    assert "libclover_mi355x.so" in ldd and "not found" not in ldd, ldd
    h = subprocess.run([str(exe), "--help"], capture_output=True, text=True, timeout=120, cwd=tmp_path)
    assert h.returncode == 0, h.stderr
    assert "--dump" in h.stdout and "Usage: cloverleaf" in h.stdout
