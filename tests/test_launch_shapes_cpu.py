"""Static check of the HIP sources in amdat_amd/csrc: every launch uses the
block size its kernel was compiled for (__launch_bounds__)."""
import glob
import os
import re

from tests.conftest import ROOT


def test_launch_block_sizes_match_bounds():
    src = "\n".join(open(f).read() for f in sorted(glob.glob(os.path.join(ROOT, "amdat_amd", "csrc", "*.hip"))))
    bounds = {}
    for m in re.finditer(r"__launch_bounds__\((\w+)(?:,[^\n]*?)?\)\s+void\s+(\w+)\(", src):
        bounds[m.group(2)] = m.group(1)
    assert len(bounds) >= 10
    launches = re.findall(r"hipLaunchKernelGGL\((\w+)(?:<[^>]*>)?,\s*[^,]+(?:\([^)]*\))?,\s*dim3\((\w+)\)", src)
    assert len(launches) >= 10
    kfn = dict(re.findall(r"auto kfn = (\w+)<", src) and [(k, k) for k in re.findall(r"auto kfn = (\w+)<", src)])
    for name, block in launches:
        if name == "kfn":
            continue
        assert name in bounds, name
        assert bounds[name] == block, (name, bounds[name], block)
    # the templated phase kernels launched through kfn use kPhaseBlock
    for k in kfn:
        assert bounds[k] == "kPhaseBlock"
    assert src.count("dim3(kPhaseBlock), lds") >= 2
