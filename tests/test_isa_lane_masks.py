"""No kernel of the library reads a VALU-compare lane mask outside the loop
that computed it (the miscompile described in isa_lane_masks.py).  Runs on
the CPU: hipcc writes the gfx950 assembly of every source."""
import os
import shutil
import tempfile
from concurrent.futures import ThreadPoolExecutor

import pytest

from conftest import ROOT, _load
from isa_lane_masks import find_hazards

# The pattern, reduced: a mask from the first loop steers a branch in the
# second one.
_BAD = """_Zkernel:
; %bb.0:
.LBB0_1:                                ; =>This Inner Loop Header: Depth=1
\tv_cmp_ne_u32_e64 s[2:3], 1, v8
\ts_cbranch_execnz .LBB0_1
; %bb.2:
.LBB0_3:                                ; =>This Inner Loop Header: Depth=1
\ts_and_b64 vcc, exec, s[2:3]
\ts_cbranch_vccnz .LBB0_3
\ts_endpgm
"""

# Safe forms: the mask computed before the loop (or in an enclosing loop)
# and read inside it, and a mask recomputed in the loop that reads it.
_GOOD = """_Zkernel:
; %bb.0:
\tv_cmp_ne_u32_e64 s[2:3], 1, v8
.LBB0_1:                                ; =>This Loop Header: Depth=1
                                        ;     Child Loop BB0_3 Depth 2
\tv_cmp_ne_u32_e64 s[4:5], 1, v9
.LBB0_3:                                ;   Parent Loop BB0_1 Depth=1
                                        ; =>  This Inner Loop Header: Depth=2
\ts_and_b64 vcc, exec, s[2:3]
\tv_cndmask_b32_e64 v1, 0, v2, s[4:5]
\tv_cmp_eq_u32_e64 s[6:7], 0, v3
\ts_andn2_b64 vcc, exec, s[6:7]
\ts_cbranch_vccnz .LBB0_3
; %bb.4:                                ;   in Loop: Header=BB0_1 Depth=1
\ts_cbranch_scc1 .LBB0_1
\ts_endpgm
"""


def test_scanner_known_answers():
    bad = find_hazards(_BAD)
    assert [(h[1], h[2], h[3]) for h in bad] == [("s[2:3]", "BB0_1", "BB0_3")]
    assert find_hazards(_GOOD) == []
    # a select after the loop is caught as well
    sel = _BAD.replace("\ts_and_b64 vcc, exec, s[2:3]", "\tv_cndmask_b32_e64 v1, 0, v2, s[2:3]")
    assert len(find_hazards(sel)) == 1


@pytest.mark.skipif(shutil.which("/opt/rocm/bin/hipcc") is None, reason="hipcc not installed")
def test_library_kernels_have_no_loop_escaping_masks():
    build = _load("ibm_build", os.path.join(ROOT, "ib.cu_amd", "build.py"))
    with tempfile.TemporaryDirectory() as tmp:
        def asm(src):
            out = os.path.join(tmp, os.path.basename(src) + ".s")
            build.device_asm(src, out)
            with open(out) as f:
                return src, find_hazards(f.read())
        with ThreadPoolExecutor(max_workers=4) as ex:
            results = list(ex.map(asm, build.sources()))
    bad = [(os.path.basename(src), h[0][:80], h[1], h[2], h[3]) for src, hz in results for h in hz]
    assert bad == [], bad
