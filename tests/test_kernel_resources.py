"""CPU check of the compiled walk kernels' resources: every variant of the grouped and
per-particle walks (and the learned inference and the upsweep) must run without a private segment
(scratch) and within the register (and, for the grouped walk, LDS) budget that keeps three
4-wave blocks per CU.  A
non-inlined helper or a spilled array shows up here as scratch bytes; the COUNT variants
once ran three times slower that way.  The device code is compiled to gfx950 assembly
(no GPU needed) and the kernel descriptors are read from its metadata."""
import os
import re
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "opengl-physics_amd", "csrc")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-w", "-Xclang", "-target-feature", "-Xclang",
         "-packed-fp32-ops", "--cuda-device-only", "-S"]


def kernel_metadata(src, tmp_path):
    out = tmp_path / (os.path.basename(src) + ".s")
    subprocess.run([HIPCC, *FLAGS, "-o", str(out), src], check=True, capture_output=True)
    text = out.read_text()
    kernels = {}
    # one metadata map per kernel: .name, .private_segment_fixed_size, .vgpr_count, ...
    for block in re.split(r"\n\s+- \.", text.split("amdhsa.kernels:")[1]):
        name = re.search(r"\.name:\s+(\S+)", block)
        if not name:
            continue
        fields = dict(re.findall(r"\.(private_segment_fixed_size|vgpr_count|group_segment_fixed_size):\s+(\d+)",
                                 block))
        kernels[name.group(1)] = {k: int(v) for k, v in fields.items()}
    return kernels


@pytest.mark.skipif(not shutil.which(HIPCC) and not os.path.exists(HIPCC), reason="hipcc not available")
@pytest.mark.parametrize("src,prefix,max_vgpr", [
    ("nbg_walk.hip", "_ZN3nbg11k_walk_bins", 168),
    ("nbg_walk_pp.hip", "_ZN3nbg9k_walk_pp", 168),
    ("nbg_learn.hip", "_ZN3nbg13k_learn_infer", 256),
    ("nbg_tree.hip", "_ZN3nbg9k_up_pair", 256),
    ("nbg_tree.hip", "_ZN3nbg8k_leaves", 256),
])
def test_walk_kernels_use_no_scratch(src, prefix, max_vgpr, tmp_path):
    kernels = kernel_metadata(os.path.join(CSRC, src), tmp_path)
    picked = {k: v for k, v in kernels.items() if k.startswith(prefix)}
    assert picked, f"no {prefix}* kernels found in {src}"
    for name, md in picked.items():
        assert md.get("private_segment_fixed_size", 0) == 0, (name, md)
        assert md["vgpr_count"] <= max_vgpr, (name, md)
        # 160 KB of LDS per CU holds three blocks of the grouped walk
        if prefix.endswith("k_walk_bins"):
            assert md["group_segment_fixed_size"] <= 160 * 1024 // 3, (name, md)
