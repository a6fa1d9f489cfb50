"""The work-queue loops of the persistent kernels as the compiler built them for gfx950 (CPU only: hipcc -S of the
kernel files). The transfer's loop pops a task, shares it through LDS between two barriers and runs it; if the
compiler's structurization puts the barriers and the ticket read in a loop that its back edge can re-enter without
the pop (lanes that take no part in the pop looping on their own), the workgroup re-reads a stale ticket forever.
Checked: in every instantiation of the transfer, the pop (the atomic on the queue head), the first barrier after it
and the LDS read of the ticket sit in the same innermost loop; in every instantiation of the element adjoints, the
pop and the stores of the block's partial sums sit in the same top-level loop (the body's end returns to the pop)."""
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "kerrsplat_amd", "csrc", "ks_k_transfer.hip")
SRC_ADJ = os.path.join(ROOT, "kerrsplat_amd", "csrc", "ks_k_adjoint.hip")
HIPCC = "/opt/rocm/bin/hipcc"


def queue_loops(asm):
    lines = asm.split("\n")
    starts = [i for i, l in enumerate(lines) if re.match(r"^_ZN3ksd17ks_chunk_transfer.*:", l)]
    out = {}
    for st in starts:
        name = lines[st].split(":")[0]
        end = next(i for i in range(st + 1, len(lines)) if lines[i].strip().startswith("s_endpgm"))
        cur, where = None, {}
        body = lines[st:end]
        for i, l in enumerate(body):
            if l.startswith(".LBB") or l.startswith("; %bb."):
                lab = l.split(":")[0].strip().lstrip(".L")
                m = re.search(r"Header=(BB\S+) Depth=(\d+)", l)
                nxt = body[i + 1] if i + 1 < len(body) else ""
                if "This Loop Header" in l or "This Loop Header" in nxt:
                    tag = l if "This Loop Header" in l else nxt
                    cur = (lab, int(re.search(r"Depth=(\d+)", tag.split("This Loop Header")[1]).group(1)))
                elif m:
                    cur = (m.group(1), int(m.group(2)))
                else:
                    cur = None
            if "global_atomic_add" in l and "pop" not in where:
                where["pop"] = cur
            elif "pop" in where and "read" not in where:
                if "s_barrier" in l and "barrier" not in where:
                    where["barrier"] = cur
                elif "barrier" in where and "ds_read_b32" in l:
                    where["read"] = cur
        out[name] = where
    return out


def compile_asm(src, out):
    flags = ["-O3", "--offload-arch=gfx950", "-std=c++17", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-fno-hip-fp32-correctly-rounded-divide-sqrt", "--cuda-device-only", "-S", "-o", str(out), src]
    subprocess.run([HIPCC] + flags, check=True, capture_output=True)
    return out.read_text()


def block_loops(asm, kernel):
    """Per instantiation of `kernel`: (the innermost loop of the first atomic add, the innermost loops of the
    64-bit global stores)."""
    lines = asm.split("\n")
    starts = [i for i, l in enumerate(lines) if re.match(rf"^_ZN3ksd{len(kernel)}{kernel}.*:", l)]
    out = {}
    for st in starts:
        end = next(i for i in range(st + 1, len(lines)) if lines[i].strip().startswith("s_endpgm"))
        body = lines[st:end]
        cur, pop, stores = None, None, set()
        for i, l in enumerate(body):
            if l.startswith(".LBB") or l.startswith("; %bb."):
                lab = l.split(":")[0].strip().lstrip(".L")
                m = re.search(r"Header=(BB\S+) Depth=(\d+)", l)
                nxt = body[i + 1] if i + 1 < len(body) else ""
                if "This Loop Header" in l or "This Loop Header" in nxt:
                    tag = l if "This Loop Header" in l else nxt
                    cur = (lab, int(re.search(r"Depth=(\d+)", tag.split("This Loop Header")[1]).group(1)))
                elif m:
                    cur = (m.group(1), int(m.group(2)))
                else:
                    cur = None
            if "global_atomic_add" in l and pop is None:
                pop = cur
            if re.match(r"\s*global_store_dwordx2", l):
                stores.add(cur)
        out[lines[st].split(":")[0]] = (pop, stores)
    return out


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc not present")
def test_element_adjoint_loop_pops_on_every_back_edge(tmp_path):
    loops = block_loops(compile_asm(SRC_ADJ, tmp_path / "adj.s"), "ks_adj_splats")
    assert len(loops) == 14          # 7 population sets x 2 precisions
    for name, (pop, stores) in loops.items():
        assert pop is not None and pop[1] == 1, (name, pop)
        assert stores and stores == {pop}, (name, pop, stores)


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc not present")
def test_queue_loop_pops_on_every_back_edge(tmp_path):
    loops = queue_loops(compile_asm(SRC, tmp_path / "tr.s"))
    assert len(loops) == 20          # 5 populations x 2 precisions x cull on/off
    for name, w in loops.items():
        assert w.get("pop") and w.get("barrier") and w.get("read"), (name, w)
        assert w["pop"] == w["barrier"] == w["read"], (name, w)
