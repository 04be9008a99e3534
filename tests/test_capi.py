"""The C-ABI library loads without a GPU and exports every entry point that
include/*.h declares; host-only plans (device = -1) build the tree, lists and
grids and refuse device work with MLFMA_ERR_STATE."""
import ctypes as C
import glob
import os
import re

import numpy as np
import pytest

from mlfma_amd import _capi as A
from mlfma_amd.plan import MlfmaError, Plan
from tests import cases

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names |= set(re.findall(r"\b(mlfma_[a-z_0-9]+)\s*\(", text))
    return names


def test_library_exports_every_declared_symbol():
    lib = A.load_library()
    names = declared_symbols()
    assert len(names) >= 14
    for n in sorted(names):
        assert hasattr(lib, n), n
    assert set(A.SIGNATURES) >= names


def test_struct_layouts_match_header(tmp_path):
    # compile the header with gcc and compare sizes and field offsets with the
    # ctypes mirror in mlfma_amd/_capi.py
    import subprocess
    structs = {"mlfma_geometry": A.Geometry, "mlfma_planewave": A.PlaneWave, "mlfma_config": A.Config,
               "mlfma_info": A.Info, "mlfma_solve_opts": A.SolveOpts, "mlfma_solve_report": A.SolveReport}
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "mlfma_c.h"', "int main(void){"]
    for cname, py in structs.items():
        lines.append(f'printf("{cname} %zu\\n", sizeof({cname}));')
        for fname, _ in py._fields_:
            lines.append(f'printf("{cname}.{fname} %zu\\n", offsetof({cname}, {fname}));')
    lines.append("return 0;}")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    out = dict(l.rsplit(" ", 1) for l in subprocess.run([str(exe)], capture_output=True, text=True,
                                                        check=True).stdout.split("\n") if l)
    for cname, py in structs.items():
        assert int(out[cname]) == C.sizeof(py), cname
        for fname, _ in py._fields_:
            assert int(out[f"{cname}.{fname}"]) == getattr(py, fname).offset, (cname, fname)


def test_host_only_plan_info_and_state_errors():
    geom, kw = cases.make("dipole_fixed")
    p = Plan(geom, cases.planewave_z(), A.make_config(device=-1, **kw))
    info = p.info()
    assert info["n"] == geom.n and info["max_level"] == 3
    assert info["nth"][3] == 9 and info["nph"][3] == 18       # a = 0.25: tau = 8
    assert info["nth"][2] == 14 and info["nph"][2] == 28      # a = 0.5:  tau = 13
    assert info["m2l_pairs"][2] > 0 and info["near_nnz"] > 0
    with pytest.raises(MlfmaError) as e:
        p.matvec(np.zeros(geom.n, complex))
    assert e.value.code == -6


def test_bad_arguments_are_errors_not_crashes():
    geom, kw = cases.make("dipole_fixed")
    lib = A.load_library()
    h = C.c_void_p()
    cfg = A.make_config(device=-1, **kw)
    cfg.interp_order = 0
    rc = lib.mlfma_plan_create(C.byref(geom.struct), C.byref(cases.planewave_z()), C.byref(cfg), C.byref(h))
    assert rc == -1 and b"interp_order" in lib.mlfma_last_error()
    pos = np.array([[0.0, 0.0, np.nan]])
    g = A.GeometryArrays(A.MLFMA_SRC_DIPOLE, pos=pos, mom=np.array([[0.0, 0.0, 1.0]]))
    rc = lib.mlfma_plan_create(C.byref(g.struct), C.byref(cases.planewave_z()),
                               C.byref(A.make_config(device=-1)), C.byref(h))
    assert rc == -5
    assert lib.mlfma_plan_create(None, None, None, None) == -1
