"""Register budget of the hot step kernels (CPU: compiles for gfx950).

The fused step is HBM-bound only while its waves neither spill much nor lose
occupancy.  The D3Q19 TRT batch kernels (the 512 x 256 x 256 throughput
workload) must keep their register target (7 waves per SIMD, at most 72
VGPRs) with at most 12 bytes of scratch per lane on rare paths (32 with the
scalar; none for FP16); a code change that makes them spill fails here (a
loop around the step body once cost 96-108 bytes per lane and 30% of the
bandwidth; 20 bytes in the FP16 kernel cost it 20%).
"""

import pathlib
import re
import shutil
import subprocess

import pytest

ROOT = pathlib.Path(__file__).resolve().parent.parent
CSRC = ROOT / "aero_amd" / "csrc"


@pytest.mark.skipif(shutil.which("hipcc") is None and not pathlib.Path("/opt/rocm/bin/hipcc").exists(),
                    reason="hipcc not available")
def test_d3q19_trt_batch_kernels_keep_their_register_budget(tmp_path):
    from aero_amd import build as B
    cmd = [B.HIPCC, *B.FLAGS, "-c", str(CSRC / "step_d3q19_trt.hip"), "-o", str(tmp_path / "t.o"),
           "-Rpass-analysis=kernel-resource-usage"]
    out = subprocess.run(cmd, capture_output=True, text=True, timeout=600)
    assert out.returncode == 0, out.stderr[-2000:]
    text = out.stderr
    blocks = re.split(r"remark: Function Name: ", text)[1:]
    seen = 0
    for b in blocks:
        name = b.split()[0]
        # batch_step_kernel<D3Q19, TRT, storage, thermal>: mangled ...ILi1ELi<ST>ELb<TH>E...
        m = re.search(r"batch_step_kernelINS_5D3Q19ELi1ELi(\d)ELb([01])ELi0E", name)
        if not m or m.group(1) == "0":  # FP64 keeps the compiler's own choice
            continue
        vgpr = int(re.search(r"VGPRs: (\d+)", b).group(1))
        scratch = int(re.search(r"ScratchSize \[bytes/lane\]: (\d+)", b).group(1))
        seen += 1
        assert vgpr <= 72, (name, vgpr)
        # without the scalar: at most 12 bytes (FP16 none); with it 32
        assert scratch <= (32 if m.group(2) == "1" else 12), (name, scratch)
        if m.groups() == ("2", "0"):
            assert scratch == 0, (name, scratch)
    assert seen == 6


@pytest.mark.skipif(shutil.which("hipcc") is None and not pathlib.Path("/opt/rocm/bin/hipcc").exists(),
                    reason="hipcc not available")
def test_d3q19_trt_local_les_kernels_keep_their_register_budget(tmp_path):
    """The local-Smagorinsky D3Q19 TRT kernels: FP32 without the scalar fits
    6 waves with no scratch; every 4- and 2-byte form stays within 80 VGPRs
    and 80 bytes of scratch per lane."""
    from aero_amd import build as B
    cmd = [B.HIPCC, *B.FLAGS, "-c", str(CSRC / "step_d3q19_trt_les.hip"), "-o", str(tmp_path / "t.o"),
           "-Rpass-analysis=kernel-resource-usage"]
    out = subprocess.run(cmd, capture_output=True, text=True, timeout=600)
    assert out.returncode == 0, out.stderr[-2000:]
    seen = 0
    for b in re.split(r"remark: Function Name: ", out.stderr)[1:]:
        m = re.search(r"batch_step_kernelINS_5D3Q19ELi1ELi(\d)ELb([01])ELi2E", b.split()[0])
        if not m or m.group(1) == "0":
            continue
        vgpr = int(re.search(r"VGPRs: (\d+)", b).group(1))
        scratch = int(re.search(r"ScratchSize \[bytes/lane\]: (\d+)", b).group(1))
        seen += 1
        assert vgpr <= 80 and scratch <= 80, (m.groups(), vgpr, scratch)
        if m.groups() == ("1", "0"):
            assert scratch == 0, (vgpr, scratch)
    assert seen == 6
