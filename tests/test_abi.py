"""The C-ABI library loads and exports every entry point include/*.h
declares; the Particle layout mirror matches a C compiler's. No GPU calls."""
import ctypes as C
import os
import re
import subprocess
import tempfile

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    for h in ("kapsel_hip.h",):
        text = open(os.path.join(ROOT, "include", h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names |= set(re.findall(r"\b(kh_[A-Za-z0-9_]+)\s*\(", text))
    return sorted(names)


def test_library_exports_every_declared_symbol():
    import kapsel3_amd
    L = kapsel3_amd.lib()
    syms = declared_symbols()
    assert len(syms) >= 20
    missing = [s for s in syms if not hasattr(L, s)]
    assert not missing, missing
    assert set(kapsel3_amd.EXPORTS) <= set(syms)


def test_strerror_without_device():
    import kapsel3_amd
    L = kapsel3_amd.lib()
    assert L.kh_strerror(0) == b"ok"
    assert L.kh_strerror(-4) == b"unsupported"


def test_particle_layout_matches_c():
    from kapsel3_amd._abi import PARTICLE_DTYPE
    src = r'''
#include <stdio.h>
#include <stddef.h>
#include "kapsel_particle.h"
int main(void) {
  printf("%zu %zu %zu %zu %zu %zu %zu %zu\n", sizeof(Particle), offsetof(Particle, x), offsetof(Particle, v),
         offsetof(Particle, omega), offsetof(Particle, q), offsetof(Particle, fr_previous),
         offsetof(Particle, surface_dw), sizeof(CTime));
  return 0;
}
'''
    with tempfile.TemporaryDirectory() as d:
        c = os.path.join(d, "p.c")
        open(c, "w").write(src)
        exe = os.path.join(d, "p")
        subprocess.run(["gcc", "-std=c11", "-I", os.path.join(ROOT, "include"), c, "-o", exe], check=True)
        vals = list(map(int, subprocess.run([exe], capture_output=True, text=True, check=True).stdout.split()))
    f = PARTICLE_DTYPE.fields
    assert vals == [PARTICLE_DTYPE.itemsize, f["x"][1], f["v"][1], f["omega"][1], f["q"][1], f["fr_previous"][1],
                    f["surface_dw"][1], 48]


def test_product_has_no_cpu_fallback():
    """The package never imports the oracle; a missing library raises."""
    pkg = os.path.join(ROOT, "kapsel3_amd")
    for fn in os.listdir(pkg):
        if fn.endswith(".py"):
            text = open(os.path.join(pkg, fn)).read()
            assert "oracle" not in re.sub(r"#.*", "", text).replace('"""', "").split("import")[0] or \
                "from oracle" not in text and "import oracle" not in text, fn
