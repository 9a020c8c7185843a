"""Kernel resources of the gfx950 build.

No kernel of the library may use private (scratch) memory.  A kernel that
writes to its by-value argument struct, or spills registers, gets a per-lane
scratch copy that every lane writes at launch; for the store's apply kernel
at a 300,000-op batch that was over a gigabyte of extra HBM writes per batch.

Two occupancies are held: the persistent engine launch needs 4 waves per
SIMD for its 1024-thread workgroups to be resident at all, and the store's
apply kernel keeps 6 (its light lists wait on dependent loads; at 4 waves per
SIMD a 512k-update batch's apply took 386 instead of 314 us).  The compiler's
resource-usage remarks name every kernel with its scratch size and occupancy.
"""
import os
import re
import subprocess
from concurrent.futures import ThreadPoolExecutor

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "risgraph_amd", "csrc")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

pytestmark = pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc not installed")


def _usage(src, out_dir):
    import sys
    sys.path.insert(0, os.path.join(ROOT, "risgraph_amd"))
    import build
    obj = os.path.join(out_dir, os.path.basename(src) + ".o")
    r = subprocess.run([HIPCC, *build.CFLAGS, "-Rpass-analysis=kernel-resource-usage", "-c", src, "-o", obj],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
    kernels, name = {}, None
    for line in r.stderr.splitlines():
        m = re.search(r"remark: Function Name: (\S+)", line)
        if m:
            name = m.group(1)
            kernels[name] = {}
            continue
        m = re.search(r"remark:\s+ScratchSize \[bytes/lane\]: (\d+)", line)
        if m and name:
            kernels[name]["scratch"] = int(m.group(1))
        m = re.search(r"remark:\s+Occupancy \[waves/SIMD\]: (\d+)", line)
        if m and name:
            kernels[name]["occupancy"] = int(m.group(1))
    return kernels


def _all_kernels(tmp_path):
    srcs = sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".hip"))
    with ThreadPoolExecutor(max_workers=4) as ex:
        results = list(ex.map(lambda s: _usage(s, str(tmp_path)), srcs))
    found = {}
    for kernels in results:
        found.update(kernels)
    return found


def test_no_kernel_uses_scratch_and_hot_kernels_keep_their_occupancy(tmp_path):
    found = _all_kernels(tmp_path)
    # the hot kernels must be among those reported
    for k in ("k_team", "k_apply_all", "k_segment", "k_expand_link", "k_scatter", "k_split"):
        assert any(k in name for name in found), k
    with_scratch = {k: v["scratch"] for k, v in found.items() if v.get("scratch")}
    assert not with_scratch, with_scratch
    occ = {k: v.get("occupancy", 0) for k, v in found.items()}
    team = [o for k, o in occ.items() if "k_team" in k]
    apply = [o for k, o in occ.items() if "k_apply_all" in k]
    assert team and min(team) >= 4, team
    assert apply and min(apply) >= 6, apply
