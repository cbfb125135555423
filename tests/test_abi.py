"""The C ABI: the HIP library exports every entry point include/fivo2d.h
declares, and the ctypes mirror has the C structs' exact sizes and offsets
(CPU only: loads nothing onto a GPU)."""
import ctypes as C
import os
import re
import subprocess
import tempfile

import pytest

import fivo2d as F
from fivo2d import _abi

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HDR = os.path.join(REPO, "include", "fivo2d.h")


def _declared():
    txt = open(HDR).read()
    return sorted(set(re.findall(r"\b(fivo_[a-z_0-9]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    assert os.path.exists(F.HIP_LIB_PATH), "libfivo2d.so not built"
    out = subprocess.run(["nm", "-D", "--defined-only", F.HIP_LIB_PATH], capture_output=True, text=True).stdout
    exported = set(l.split()[-1] for l in out.splitlines() if l.strip())
    missing = [s for s in _declared() if s not in exported]
    assert not missing, missing
    assert len(_declared()) >= 25


def test_struct_layout_matches_header():
    src = r'''
#include <stdio.h>
#include <stddef.h>
#include "fivo2d.h"
int main(void) {
  printf("%zu %zu %zu %zu %zu %zu %zu\n", sizeof(fivo_grid_desc), sizeof(fivo_eos_desc),
         sizeof(fivo_model_desc), sizeof(fivo_work_view), sizeof(fivo_run_opts),
         sizeof(fivo_run_result), sizeof(fivo_counts));
  printf("%zu %zu %zu %zu\n", offsetof(fivo_model_desc, terms), offsetof(fivo_model_desc, eta_over_s),
         offsetof(fivo_model_desc, E_vac_cut), offsetof(fivo_run_opts, on_dump));
  return 0;
}
'''
    with tempfile.TemporaryDirectory() as d:
        c = os.path.join(d, "t.c")
        open(c, "w").write(src)
        exe = os.path.join(d, "t")
        subprocess.run(["gcc", "-I", os.path.join(REPO, "include"), c, "-o", exe], check=True)
        lines = subprocess.run([exe], capture_output=True, text=True, check=True).stdout.split("\n")
    sizes = [int(v) for v in lines[0].split()]
    offs = [int(v) for v in lines[1].split()]
    assert sizes == [C.sizeof(_abi.GridDesc), C.sizeof(_abi.EosDesc), C.sizeof(_abi.ModelDesc),
                     C.sizeof(_abi.WorkView), C.sizeof(_abi.RunOpts), C.sizeof(_abi.RunResult),
                     C.sizeof(_abi.Counts)]
    assert offs == [_abi.ModelDesc.terms.offset, _abi.ModelDesc.eta_over_s.offset,
                    _abi.ModelDesc.E_vac_cut.offset, _abi.RunOpts.on_dump.offset]


def test_refused_knobs():
    with pytest.raises(ValueError):
        F.build_model_2d(diff_dt_coeff=0.5)
    with pytest.raises(ValueError, match="Pick ONE"):
        F.build_model_2d(advect_nu=True)                          # relax_advect_nu defaults to true
    F.build_model_2d(advect_nu=True, relax_advect_nu=False)       # flux transport alone is built
    with pytest.raises(ValueError):
        F.build_model_2d(enable_diff=True, consistent_fm=True, tauN_coeff=2.0)
