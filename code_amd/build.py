"""Build recipe: the HIP library for gfx950, the Fortran binding, and the
oracle's C restatement (test infrastructure).

All outputs go to in-tree ``_build/`` folders so they travel with the tree.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent
PKG = ROOT / "code_amd"
CSRC = PKG / "csrc"
BUILD = PKG / "_build"
ORACLE = ROOT / "oracle"
ORACLE_BUILD = ORACLE / "_build"

LIB_NAME = "libgetm_transport.so"
ORACLE_LIB = "libgetm_oracle.so"

HIP_SOURCES = ["gt_kernels.hip", "gt_api.hip", "gt_batched_f64.hip", "gt_batched_f32_g1.hip",
               "gt_batched_f32_g3.hip", "gt_batched_f32_g6.hip", "gt_batched_f32_g9.hip",
               "gt_adjoint.hip", "gt_substeps.hip", "gt_eqstate.hip", "gt_inputs.hip",
               "gt_diffusion.hip", "gt_momentum.hip", "gt_restart.cpp", "gt_ip_fused.hip"]

# -ffp-contract=off: no fused multiply-add, so FP64 results round exactly as
# the CPU build's.  Packed FP32 instructions are disabled on purpose.
HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def _run(cmd, cwd=None):
    r = subprocess.run(cmd, cwd=cwd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(map(str, cmd)) + "\n" + r.stdout + r.stderr)
        raise RuntimeError(f"build step failed: {cmd[0]} (exit {r.returncode})")
    return r


def hipcc() -> str:
    p = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    return p


def build_hip(verbose=False) -> Path:
    BUILD.mkdir(parents=True, exist_ok=True)
    srcs = [CSRC / s for s in HIP_SOURCES if (CSRC / s).exists()]
    objs = [BUILD / (s.stem + ".o") for s in srcs]

    def comp(pair):
        s, o = pair
        _run([hipcc(), *HIPCC_FLAGS, "-c", str(s), "-o", str(o)])
        return o

    with ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        list(ex.map(comp, zip(srcs, objs)))
    out = BUILD / LIB_NAME
    tmp = BUILD / (LIB_NAME + ".tmp")
    _run([hipcc(), "--offload-arch=gfx950", "-shared", "-fPIC", *map(str, objs), "-o", str(tmp)])
    os.replace(tmp, out)
    return out


def build_fortran() -> Path | None:
    """Compile the ISO_C_BINDING module (proves the Fortran surface builds)
    and the Fortran driver used by the Fortran parity test."""
    flang = shutil.which("flang") or "/opt/rocm/lib/llvm/bin/flang"
    src = PKG / "fortran" / "getm_transport_gpu.F90"
    if not src.exists() or not Path(flang).exists():
        return None
    fb = BUILD / "fortran"
    fb.mkdir(parents=True, exist_ok=True)
    obj = fb / "getm_transport_gpu.o"
    _run([flang, "-cpp", "-O2", "-fPIC", "-c", str(src), "-o", str(obj),
          "-module-dir", str(fb)])
    link = ["-L", str(BUILD), "-lgetm_transport", f"-Wl,-rpath,{BUILD}",
            "-Wl,-rpath,$ORIGIN/.."]
    for name in ("adv_driver", "rows_driver"):
        drv = PKG / "fortran" / f"{name}.F90"
        if drv.exists():
            exe = fb / name
            _run([flang, "-cpp", "-O2", str(drv), str(obj), "-I", str(fb),
                  "-module-dir", str(fb), *link, "-o", str(exe)])
    # the external routines with the reference's argument lists, built
    # against the test's stand-ins of the host model's modules
    ext = [PKG / "fortran" / n for n in ("ext_host_modules.F90", "getm_external.F90",
                                         "ext_driver.F90")]
    if all(p.exists() for p in ext):
        eb = fb / "ext"
        eb.mkdir(parents=True, exist_ok=True)
        objs = []
        for src in ext[:2]:
            o = eb / (src.stem + ".o")
            _run([flang, "-cpp", "-O2", "-c", str(src), "-I", str(fb), "-I", str(eb),
                  "-module-dir", str(eb), "-o", str(o)])
            objs.append(str(o))
        _run([flang, "-cpp", "-O2", str(ext[2]), *objs, str(obj), "-I", str(fb), "-I", str(eb),
              "-module-dir", str(eb), *link, "-o", str(fb / "ext_driver")])
    return obj


def build_oracle() -> Path:
    """The oracle's C restatement (test infrastructure, never shipped)."""
    ORACLE_BUILD.mkdir(parents=True, exist_ok=True)
    out = ORACLE_BUILD / ORACLE_LIB
    tmp = ORACLE_BUILD / (ORACLE_LIB + ".tmp")
    _run(["gcc", "-O2", "-std=c11", "-ffp-contract=off", "-fno-fast-math", "-fopenmp",
          "-fPIC", "-shared", str(ORACLE / "getm_oracle.c"), "-o", str(tmp), "-lm"])
    os.replace(tmp, out)
    return out


def build_all(verbose=False):
    build_oracle()
    lib = build_hip(verbose)
    build_fortran()
    return lib


if __name__ == "__main__":
    print(build_all(verbose=True))
