"""Structure of the persistent work-queue kernels in the gfx950 code the compiler emits.

A persistent kernel's workgroup takes items in a loop: thread 0 bumps the queue head, a barrier
publishes the item, a second barrier protects the slot.  If the compiler's control-flow
structurizer moves thread 0's queue access out of the loop that holds those barriers (it did once,
when thread 0 had a second block at the bottom of the loop), the other threads spin over the
barriers on a stale item and the grid never drains.  This test compiles trace.hip to assembly (hipcc,
no GPU needed) and checks, for every instantiation of the persistent primary traversal of batched
updates, that the queue atomic and the first barriers after it sit in the same outermost loop."""
import os
import re
import shutil
import subprocess

import pytest

CSRC = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "vulkanexp_amd", "csrc")
HIPCC = "/opt/rocm/bin/hipcc"
FLAGS = ["-O3", "-std=c++17", "--offload-arch=gfx950", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-I" + os.path.join(os.path.dirname(os.path.dirname(CSRC)), "include"), "--cuda-device-only", "-S"]

pytestmark = pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc not available")


def kernel_events(asm: str, name_re: str):
    """{kernel: [(event, loop depth)]} for queue atomics ('A') and barriers ('B')."""
    out, cur, depth = {}, None, 0
    for ln in asm.split("\n"):
        m = re.match(r"^(_Z\w*(?:" + name_re + r")\w*):", ln)
        if m:
            cur, depth = m.group(1), 0
            out[cur] = []
            continue
        if cur is None:
            continue
        if ln.startswith(".LBB") or ln.startswith("; %bb") or "Loop Header" in ln:
            hdr = re.search(r"This (?:Inner )?Loop Header: Depth=(\d+)", ln)
            inner = re.search(r"Header=BB\d+_\d+ Depth=(\d+)", ln)
            if hdr:
                depth = int(hdr.group(1))
            elif inner:
                depth = int(inner.group(1))
            elif ln.startswith(".LBB") or ln.startswith("; %bb"):
                depth = 0
        if re.search(r"\bglobal_atomic_add\b", ln) and "s[" in ln:
            out[cur].append(("A", depth))
        if "s_barrier" in ln:
            out[cur].append(("B", depth))
        if "s_endpgm" in ln:
            cur = None
    return out


@pytest.fixture(scope="module")
def asm(tmp_path_factory):
    d = tmp_path_factory.mktemp("asm")
    texts = {}
    for src, contract in (("trace.hip", "-ffp-contract=off"),):
        out = d / (src + ".s")
        subprocess.run([HIPCC, *FLAGS, contract, os.path.join(CSRC, src), "-o", str(out)], check=True,
                       capture_output=True, timeout=600)
        texts[src] = out.read_text()
    return texts


def test_persistent_kernels_fetch_inside_the_barrier_loop(asm):
    kernels = {}
    kernels.update(kernel_events(asm["trace.hip"], "k_batch_trace"))
    assert len(kernels) >= 3, sorted(kernels)  # 3 workgroup sizes
    for name, ev in kernels.items():
        atomics = [i for i, (e, _) in enumerate(ev) if e == "A"]
        assert atomics, f"{name}: no queue atomic"
        first = atomics[0]
        depth = ev[first][1]
        assert depth == 1, f"{name}: queue atomic at loop depth {depth}"
        barriers = [d for e, d in ev[first + 1:] if e == "B"][:2]
        # single-wave workgroups have their barriers removed; otherwise both item barriers follow the
        # atomic in the same outermost loop
        if any(e == "B" for e, _ in ev):
            assert barriers == [1, 1], f"{name}: item barriers at depths {barriers}"
