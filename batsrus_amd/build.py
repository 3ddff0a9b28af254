"""Build step for the native parts: the gfx950 HIP library (C ABI) and the
Fortran ISO_C_BINDING bridge.  Everything is built in-tree so the shared
objects travel with the source tree."""
from __future__ import annotations

import os
import shutil
import subprocess
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLANG = os.environ.get("FC_AMD", "/opt/rocm/bin/amdflang")

# gfx950 only; packed FP32 instructions are disabled for all device code.
HIP_FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wno-unused-command-line-argument"]

# the fused stage kernel, one object per (block size, face flux, storage
# type)
STAGE_UNITS = [(nb, flux, real) for nb in (8, 16)
               for flux in ("rusanov", "linde", "hlld")
               for real in ("f64", "f32")]


def _run(cmd, cwd=ROOT):
    r = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE,
                       stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("build step failed: %s\n%s" % (" ".join(cmd),
                                                          r.stdout[-4000:]))
    return r.stdout


def build_hip(jobs=None):
    """Compile every HIP translation unit for gfx950 (in parallel) and link
    batsrus_amd/libbats_hip.so."""
    csrc = os.path.join(HERE, "csrc")
    odir = os.path.join(csrc, "_build")
    os.makedirs(odir, exist_ok=True)
    inc = ["-I" + os.path.join(ROOT, "include"), "-I" + csrc]
    units = [(os.path.join(csrc, name + ".hip"),
              os.path.join(odir, name + ".o"), [])
             for name in ("bats_hip", "mlp", "surrogate")]
    for nb, flux, real in STAGE_UNITS:
        units.append((os.path.join(csrc, "stage_inst.hip"),
                      os.path.join(odir, "stage_%d_%s_%s.o" % (nb, flux, real)),
                      ["-DBATS_NB=%d" % nb, "-DBATS_FLUX=%s" % flux,
                       "-DBATS_REAL=%s" % real]))
    jobs = jobs or min(len(units), os.cpu_count() or 1, 16)
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        futs = [ex.submit(_run, [HIPCC] + HIP_FLAGS + inc + defs +
                          ["-c", src, "-o", obj])
                for src, obj, defs in units]
        for f in futs:
            f.result()
    out = os.path.join(HERE, "libbats_hip.so")
    tmp = out + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] +
         [obj for _, obj, _ in units])
    os.replace(tmp, out)
    return out


def build_fortran():
    """Compile the bind(C) bridge module and the demonstration driver
    against libbats_hip.so with AMD flang, when it is installed."""
    if not (os.path.exists(FLANG) or shutil.which("amdflang")):
        return None
    fc = FLANG if os.path.exists(FLANG) else shutil.which("amdflang")
    fdir = os.path.join(HERE, "fortran")
    bdir = os.path.join(fdir, "_build")
    os.makedirs(bdir, exist_ok=True)
    exe = os.path.join(fdir, "bats_fortran_driver")
    srcs = [os.path.join(fdir, "ModHipBridge.f90"),
            os.path.join(fdir, "bats_fortran_driver.f90")]
    if not all(os.path.exists(s) for s in srcs):
        return None
    _run([fc, "-O2", "-cpp", "-fdefault-real-8", "-J", bdir, "-o", exe] +
         srcs + ["-L" + HERE, "-lbats_hip", "-Wl,-rpath,$ORIGIN/.."])
    return exe


def build_all():
    built = [build_hip()]
    f = build_fortran()
    if f:
        built.append(f)
    return built
