"""Register and scratch use of the traversal kernels, read from the gfx950 assembly hipcc writes.

A per-lane array that the compiler cannot keep in registers (one indexed with a value known only
at run time, including a select chain it folds back into an index) is placed in scratch memory:
every access then goes through the memory hierarchy and the kernel slows down with no change in
its results, so the parity tests cannot see it.  The default-occupancy variants of the ray
traversal kernels must use no scratch and no spilled registers.  Needs only hipcc (no device).
"""
import os
import re
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "dynamiclighting-bevy_amd", "csrc", "dl_kernels.hip")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

# mangled-name prefixes: the closest-hit and any-hit ray streams without the 8-wave register cap,
# the per-lane shadow trace, the per-texel burial kernel and the photon kernels
KERNELS = {
    "k_any_stream<1, closest>": "_ZN2dl12k_any_streamILi1ELb1E",
    "k_any_stream<1, any>": "_ZN2dl12k_any_streamILi1ELb0E",
    "k_trace<false>": "_ZN2dl7k_traceILb0E",
    "k_burial_lanes": "_ZN2dl14k_burial_lanes",
    "k_photons": "_ZN2dl9k_photons",
    "k_photons_listed": "_ZN2dl16k_photons_listed",
    "k_gather_split": "_ZN2dl14k_gather_split",
    "k_gather_open": "_ZN2dl13k_gather_open",
}


@pytest.fixture(scope="module", params=[4, 8], ids=["bvh4", "bvh8"])
def kernel_meta(request, tmp_path_factory):
    """Kernel metadata of the default build (4-wide BVH nodes) and of the DL_BVH_WIDTH=8 build."""
    if not (os.path.exists(HIPCC) or shutil.which(HIPCC)):
        pytest.skip("hipcc not found")
    out = str(tmp_path_factory.mktemp("isa") / "dl_kernels.s")
    cmd = [HIPCC, "--offload-arch=gfx950", "-O3", "-std=c++17", "-ffp-contract=off", "-Xclang", "-target-feature",
           "-Xclang", "-packed-fp32-ops", "-Wno-unused-command-line-argument", f"-DDL_BVH_WIDTH={request.param}",
           "--cuda-device-only", "-S", SRC, "-o", out]
    subprocess.run(cmd, check=True, capture_output=True, timeout=600)
    text = open(out).read()
    meta = {}
    # the code object metadata lists each kernel as a YAML record starting at `- .agpr_count:`
    for rec in text.split("  - .agpr_count:")[1:]:
        name = re.search(r"\.name:\s+(\S+)", rec).group(1)
        meta[name] = {
            "scratch": int(re.search(r"\.private_segment_fixed_size:\s+(\d+)", rec).group(1)),
            "spills": int(re.search(r"\.vgpr_spill_count:\s+(\d+)", rec).group(1)),
            "vgprs": int(re.search(r"\.vgpr_count:\s+(\d+)", rec).group(1)),
        }
    return meta


@pytest.mark.parametrize("label", sorted(KERNELS))
def test_traversal_kernels_use_no_scratch(kernel_meta, label):
    prefix = KERNELS[label]
    found = {n: m for n, m in kernel_meta.items() if n.startswith(prefix)}
    assert found, f"{label}: no kernel named {prefix}* in the assembly"
    for name, m in found.items():
        assert m["scratch"] == 0 and m["spills"] == 0, f"{label} ({name[:60]}): {m}"
