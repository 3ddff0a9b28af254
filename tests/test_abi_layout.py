"""The ctypes mirrors of the C structs match the C layout: every field of
batsrus_amd.core.BatsConfig (include/bats_hip.h) and pyoracle.ObConfig
(oracle/bats_oracle.h) sits at the offset gcc gives it, and the sizes
agree, so a field added on one side only fails here instead of silently
shifting the rest."""
import ctypes as C
import os
import shutil
import subprocess
import tempfile

import pytest

import pyoracle as po
from batsrus_amd.core import BatsConfig

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _c_layout(header, ctype, fields):
    src = ['#include <stddef.h>', '#include <stdio.h>',
           '#include "%s"' % header, 'int main(void) {',
           '  printf("size %%zu\\n", sizeof(%s));' % ctype]
    for name in fields:
        src.append('  printf("%s %%zu\\n", offsetof(%s, %s));'
                   % (name, ctype, name))
    src += ['  return 0;', '}']
    with tempfile.TemporaryDirectory() as d:
        c = os.path.join(d, "layout.c")
        exe = os.path.join(d, "layout")
        with open(c, "w") as f:
            f.write("\n".join(src) + "\n")
        subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), "-I",
                        os.path.join(ROOT, "oracle"), c, "-o", exe],
                       check=True, capture_output=True)
        out = subprocess.run([exe], check=True, capture_output=True,
                             text=True).stdout
    return {k: int(v) for k, v in (l.split() for l in out.splitlines())}


@pytest.mark.skipif(shutil.which("gcc") is None, reason="gcc not found")
@pytest.mark.parametrize("header,ctype,cls", [
    ("bats_hip.h", "bats_config", BatsConfig),
    ("bats_oracle.h", "ob_config", po.ObConfig),
])
def test_ctypes_mirror_matches_c(header, ctype, cls):
    names = [f[0] for f in cls._fields_]
    lay = _c_layout(header, ctype, names)
    assert lay["size"] == C.sizeof(cls)
    for name in names:
        assert lay[name] == getattr(cls, name).offset, name


def _fortran_fields():
    src = open(os.path.join(ROOT, "batsrus_amd", "fortran",
                            "ModHipBridge.f90")).read().splitlines()
    i0 = next(i for i, l in enumerate(src)
              if l.strip().lower().startswith("type, bind(c)")
              and l.strip().lower().endswith("bats_config"))
    names = []
    for line in src[i0 + 1:]:
        line = line.split("!")[0].strip()
        if line.lower().startswith("end type"):
            break
        if "::" not in line:
            continue
        for item in line.split("::", 1)[1].split(","):
            names.append(item.split("=")[0].strip())
    return names


_FC = os.environ.get("FC_AMD", "/opt/rocm/bin/amdflang")
_LIB = os.path.join(ROOT, "batsrus_amd", "libbats_hip.so")


@pytest.mark.skipif(not (os.path.exists(_FC) and os.path.exists(_LIB)
                         and shutil.which("gcc")),
                    reason="amdflang, gcc or the built library missing")
def test_fortran_type_matches_c():
    """type(bats_config) of the ISO_C_BINDING bridge has the C struct's
    field order, offsets and size (c_loc offsets, c_sizeof)."""
    names = _fortran_fields()
    lay = _c_layout("bats_hip.h", "bats_config", names)
    prog = ["program layout", "  use iso_c_binding", "  use ModHipBridge",
            "  implicit none", "  type(bats_config), target:: c",
            "  integer(c_intptr_t):: b",
            "  b = transfer(c_loc(c), b)",
            "  print '(a,1x,i0)', 'size', c_sizeof(c)"]
    for n in names:
        prog.append("  print '(a,1x,i0)', '%s', transfer(c_loc(c%%%s), b) - b"
                    % (n, n))
    prog.append("end program layout")
    with tempfile.TemporaryDirectory() as d:
        f = os.path.join(d, "layout.f90")
        exe = os.path.join(d, "layout")
        with open(f, "w") as fh:
            fh.write("\n".join(prog) + "\n")
        subprocess.run([_FC, "-cpp", "-J", d, "-o", exe,
                        os.path.join(ROOT, "batsrus_amd", "fortran",
                                     "ModHipBridge.f90"), f,
                        "-L" + os.path.dirname(_LIB), "-lbats_hip",
                        "-Wl,-rpath," + os.path.dirname(_LIB)],
                       check=True, capture_output=True)
        out = subprocess.run([exe], check=True, capture_output=True,
                             text=True).stdout
    got = {k: int(v) for k, v in (l.split() for l in out.splitlines())}
    assert got["size"] == lay["size"]
    for n in names:
        assert got[n] == lay[n], n
