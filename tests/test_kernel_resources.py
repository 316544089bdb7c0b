"""Register budget of the entry-parallel stack kernels, on 16x16 tiles (smm_mfma_ep) and on 4x4 blocks (smm_mfma_eb)
(CPU, hipcc only). Kernels of fewer than 16 output tiles (block instructions)
keep a window of the stack in registers and read it across lanes (readlane, shuffles); a spilled window
register reloaded under a partial EXEC mask would hand those reads stale lanes, so every such kernel must
compile without spills. The 16-tile kernels keep loading their stack entries and may spill (the k-split 4x4-block
kernels, smm_mfma_ek, keep the window up to 16 block instructions)."""
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "libxstream_amd", "csrc")
HIPCC = os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "bin", "hipcc")


@pytest.mark.parametrize("real,tag", [("double", "f64"), ("float", "f32")])
def test_windowed_ep_kernels_do_not_spill(tmp_path, real, tag):
    if not os.access(HIPCC, os.X_OK):
        pytest.skip("hipcc not available")
    out = subprocess.run([HIPCC, "--offload-arch=gfx950", "-O3", "-std=c++17", "-munsafe-fp-atomics", "-Xclang",
                          "-target-feature", "-Xclang", "-packed-fp32-ops", "-I" + os.path.join(ROOT, "include"),
                          "-I" + SRC, "-DLXS_REAL=" + real, "-DLXS_TAG=" + tag, "--cuda-device-only", "-c",
                          os.path.join(SRC, "smm_mfma.hip"), "-o", str(tmp_path / "k.o"),
                          "-Rpass-analysis=kernel-resource-usage"], capture_output=True, text=True, timeout=600)
    assert out.returncode == 0, out.stderr[-2000:]
    kernels, name = {}, None
    for line in out.stderr.splitlines():
        m = re.search(r"Function Name: (\S+)", line)
        if m:
            name = m.group(1)
            continue
        m = re.search(r"(VGPRs Spill|ScratchSize \[bytes/lane\]): (\d+)", line)
        if m and name is not None and ("smm_mfma_ep" in name or "smm_mfma_eb" in name or "smm_mfma_ek" in name):
            kernels.setdefault(name, {})[m.group(1)] = int(m.group(2))
    checked = 0
    for name, res in kernels.items():
        ti, tj = (int(x) for x in re.search(r"smm_mfma_e[pbk]I[df]Li(\d+)ELi(\d+)E", name).groups())
        if ti * tj < 16 or ("smm_mfma_ek" in name and ti * tj <= 16):
            assert res.get("VGPRs Spill", 1) == 0 and res.get("ScratchSize [bytes/lane]", 1) == 0, (name, res)
            checked += 1
    assert checked >= 10, kernels
