"""Compiler check on the gfx950 code of the product kernels (CPU only: hipcc emits the assembly).

A divergent loop whose exit condition is tested again after the loop must carry that condition per
lane across iterations.  hipcc has been seen to reuse the compare mask of the loop's last
iteration instead, in which the bits of lanes that left the loop earlier are zero; in the capture
pass this sent nodes to the wrong kernel.  This test compiles engine.hip to assembly and fails if
a block that exits a loop reads, before redefining it, a lane mask (vcc or an SGPR pair) whose
last definition is a vector compare inside that loop."""
import os
import re
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


def _sregs(tok):
    m = re.fullmatch(r"s\[(\d+):(\d+)\]", tok)
    if m:
        return set(range(int(m.group(1)), int(m.group(2)) + 1))
    return {"vcc"} if tok == "vcc" else set()


def _operands(ins):
    op = ins.split()[0]
    return op, [a.strip() for a in ins[len(op):].split(",")]


def _defs(op, args):
    if op.startswith("v_cmp") and op.endswith("_e32"):
        return {"vcc"}
    if (op.startswith("v_div_scale") or re.match(r"v_\w+_co_\w+", op)) and len(args) > 1:
        return _sregs(args[0]) | _sregs(args[1])
    return _sregs(args[0]) if args and args[0] else set()


def stale_loop_exit_masks(asm_lines, prefix="_ZN3sbr"):
    """(kernel, block, instruction, defining compare) for each suspicious read."""
    out = []
    func, prev_depth, headers = None, 0, []
    for i, line in enumerate(asm_lines):
        m = re.match(r"^(_Z\S+):", line)
        if m:
            func = m.group(1) if m.group(1).startswith(prefix) else None
            prev_depth, headers = 0, []
        if func is None or not re.match(r"^(\.LBB\d+_\d+|; %bb\.\d+)", line):
            continue
        md = re.search(r"Depth=(\d+)", line)
        depth = int(md.group(1)) if md else 0
        nxt = asm_lines[i + 1] if i + 1 < len(asm_lines) else ""
        mh = re.search(r"This (Inner )?Loop Header: Depth=(\d+)", nxt)
        if mh:
            depth = int(mh.group(2))
            headers.append((i, depth))
        if depth < prev_depth:
            hdr = max([h for h in headers if h[1] == prev_depth and h[0] < i], default=None)
            written = set()
            for k in range(i + 1, min(i + 40, len(asm_lines))) if hdr else ():
                ins = asm_lines[k].strip()
                if not ins or ins.startswith(";"):
                    continue
                if ins.startswith((".LBB", "s_branch", "s_cbranch")):
                    break
                op, args = _operands(ins)
                if op.startswith("v_div_scale") or re.match(r"v_\w+_co_\w+", op):
                    written |= _defs(op, args)
                    continue
                srcs = set().union(*[_sregs(a) for a in args[1:]]) if len(args) > 1 else set()
                if op == "v_cndmask_b32_e32":
                    srcs.add("vcc")
                for r in srcs - written:
                    for q in range(i - 1, hdr[0], -1):   # last definition inside the loop
                        x = asm_lines[q].strip()
                        if not x or x.startswith(";") or x.startswith("."):
                            continue
                        xo, xa = _operands(x)
                        if r in _defs(xo, xa):
                            if xo.startswith("v_cmp"):
                                out.append((func, asm_lines[i].split()[0], ins, x))
                            break
                written |= _defs(op, args)
        prev_depth = depth
    return out


@pytest.mark.skipif(not shutil.which(HIPCC) and not os.path.exists(HIPCC), reason="hipcc not installed")
def test_no_loop_exit_reads_a_last_iteration_mask(tmp_path):
    asm = tmp_path / "engine.s"
    subprocess.check_call([HIPCC, "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
                           "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-Wno-unused-result",
                           "-I", os.path.join(ROOT, "include"), "-I", os.path.join(ROOT, "sbrsolver_amd", "csrc"),
                           "-x", "hip", "--cuda-device-only", "-S",
                           os.path.join(ROOT, "sbrsolver_amd", "csrc", "engine.hip"), "-o", str(asm)],
                          stderr=subprocess.DEVNULL)
    lines = asm.read_text().split("\n")
    assert sum(1 for l in lines if re.match(r"^_ZN3sbr\S*k_trace\S*:", l)) >= 1
    assert stale_loop_exit_masks(lines) == []


def test_scanner_flags_a_last_iteration_mask():
    """The pattern as the compiler emitted it for the capture pass's counting loop."""
    asm = """_ZN3sbr9k_demoEv:
.LBB0_1:                                ;   in Loop: Header=BB0_0 Depth=1
	s_mov_b64 s[12:13], 0
.LBB0_2:                                ;   Parent Loop BB0_0 Depth=1
                                        ; =>  This Inner Loop Header: Depth=2
	v_add_u32_e32 v0, v20, v0
	v_cmp_ge_i32_e64 s[6:7], v29, v98
	v_cmp_lt_i32_e32 vcc, 12, v0
	s_or_b64 s[6:7], s[6:7], vcc
	s_and_b64 s[6:7], exec, s[6:7]
	s_or_b64 s[12:13], s[6:7], s[12:13]
	s_andn2_b64 exec, exec, s[12:13]
	s_cbranch_execnz .LBB0_2
.LBB0_3:                                ;   in Loop: Header=BB0_0 Depth=1
	s_or_b64 exec, exec, s[12:13]
	s_and_b64 s[6:7], vcc, exec
""".split("\n")
    hits = stale_loop_exit_masks(asm)
    assert len(hits) == 1 and "vcc" in hits[0][2]
    fixed = [l.replace("s_and_b64 s[6:7], vcc, exec", "s_orn2_b64 s[6:7], s[16:17], exec") for l in asm]
    assert stale_loop_exit_masks(fixed) == []
