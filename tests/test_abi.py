"""C-ABI checks that need no GPU: the library loads, exports every symbol
include/bm4dof.h declares, and its host-side parameter logic equals the
oracle's restatement of initializeNlbParameters / set* (nldct.cpp:81-210)
and of the bm4d-of command line's override order (main_vnlb.cpp:206-246)."""
import ctypes as C
import os
import re

import pytest

import oracle as O
from conftest import REPO, bm4dof_amd as A


def _declared_functions():
    text = open(os.path.join(REPO, "include", "bm4dof.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?[a-zA-Z_][\w\s\*]*?\b([a-zA-Z_]\w*)\s*\(", text, flags=re.M)
    return sorted({n for n in names if n not in ("if", "while", "for", "return", "sizeof")})


def test_library_loads_and_exports_header():
    L = A.lib()
    declared = _declared_functions()
    assert "bm4dof_run" in declared and "Dual_TVL1_optic_flow_multiscale" in declared
    missing = [n for n in declared if not hasattr(L, n)]
    assert not missing, missing
    assert sorted(A.EXPORTS) == declared


def test_params_struct_layout():
    assert C.sizeof(A.Params) == 23 * 4
    assert [f for f, _ in A.Params._fields_] == [f for f, _ in O.Params._fields_]


CASES = [
    # (step, sigma, frames, cli overrides)
    (1, 20, 20, dict(px=4, pt=4, wx=11, np_=32)),
    (2, 20, 20, dict(px=4, pt=4, wx=11, np_=32)),
    (1, 20, 1, dict(px=8, pt=1, wx=21, np_=16, beta=3.0)),
    (2, 20, 1, dict(px=8, pt=1, wx=21, np_=16)),
    (1, 45, 1, dict()),
    (2, 45, 1, dict()),
    (1, 40, 30, dict(px=4, pt=4, wx=15, np_=48)),
    (2, 10, 5, dict(px=7, wx=5, np_=1000)),
    (1, 30, 9, dict(px=5, pt=2, np_=4, beta=2.0, beta_mean=1.5, order_inv=True, no_step=True)),
]


@pytest.mark.parametrize("step,sigma,frames,kw", CASES)
def test_params_match_oracle(step, sigma, frames, kw):
    a = A.cli_params(step, sigma, (frames, 1, 64, 64), **kw)
    o = O.cli_params(step, sigma, frames, **kw)
    assert a.as_dict() == o.as_dict()


def test_reference_defaults_video_and_image():
    p1 = O.init_params(1, 20, 20)
    p2 = O.init_params(2, 20, 20)
    assert (p1.sizePatch, p1.nSimilarPatches, p1.sizeSearchWindow, p1.offSet) == (8, 8, 7, 6)
    assert (p2.sizePatch, p2.offSet, p2.transform) == (7, 4, 0)
    assert abs(p1.tau - 67.1) < 1e-5 and abs(p2.tau - 54.8) < 1e-5
    q = O.cli_params(1, 20, 20, px=4, pt=4, wx=11, np_=32)
    assert (q.boundary, q.offSet, q.transform, q.nSimilarPatches) == (13, 3, 0, 32)
    q = O.cli_params(2, 20, 20, px=4, pt=4, wx=11, np_=32)
    assert (q.boundary, q.offSet) == (13, 2)
    i1 = O.init_params(1, 20, 1)
    assert (i1.sizePatch, i1.nSimilarPatches, i1.sizeSearchWindow, i1.transform) == (8, 16, 39, 1)
