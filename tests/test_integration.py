"""The reference-side bindings in integration/ compile against the
reference's own headers (higtree/src, higflow/src) with -Wall -Wextra
-Werror, in 2D and 3D, and every library entry point they call is declared
in include/higflow_hip.h and exported by the built library.  Skipped where
the reference tree or the MPI/glib headers its headers include are absent."""
import os
import re
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = "/root/reference"
CONDA = "/opt/conda"
HDRS = [os.path.join(REF, "higtree", "src"), os.path.join(REF, "higflow", "src"),
        os.path.join(CONDA, "include"), os.path.join(CONDA, "include", "glib-2.0"),
        os.path.join(CONDA, "lib", "glib-2.0", "include")]

needs = pytest.mark.skipif(not (os.path.isdir(REF) and os.path.exists(os.path.join(CONDA, "include", "mpi.h"))
                                and shutil.which("gcc")),
                           reason="reference headers or the MPI/glib headers they include are absent")


def compile_obj(src, dim, tmp_path):
    out = str(tmp_path / ("%s_%dd.o" % (os.path.basename(src)[:-2], dim)))
    cmd = ["gcc", "-std=gnu11", "-Wall", "-Wextra", "-Werror", "-c", "-DDIM=%d" % dim,
           "-I" + os.path.join(ROOT, "include"), "-I" + os.path.join(ROOT, "integration")]
    # the reference's headers carry unused static definitions of their own:
    # include them as system headers so only this code's warnings count
    for h in HDRS:
        cmd += ["-isystem", h]
    r = subprocess.run(cmd + [os.path.join(ROOT, "integration", src), "-o", out], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return out


def undefined_hfh(obj):
    r = subprocess.run(["nm", "-u", obj], capture_output=True, text=True, check=True)
    return sorted({ln.split()[-1] for ln in r.stdout.splitlines() if ln.split()[-1].startswith("hfh_")})


@needs
@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("src", ["solver-hip.c", "higflow-step-hip.c"])
def test_binding_compiles_against_reference_headers(src, dim, tmp_path):
    obj = compile_obj(src, dim, tmp_path)
    used = undefined_hfh(obj)
    assert used
    header = open(os.path.join(ROOT, "include", "higflow_hip.h")).read()
    declared = set(re.findall(r"\b(hfh_\w+)\s*\(", header))
    assert set(used) <= declared, set(used) - declared
    lib = os.path.join(ROOT, "higflow_amd", "libhigflow_hip.so")
    if os.path.exists(lib):
        r = subprocess.run(["nm", "-D", "--defined-only", lib], capture_output=True, text=True, check=True)
        exported = {ln.split()[-1] for ln in r.stdout.splitlines()}
        assert set(used) <= exported, set(used) - exported
