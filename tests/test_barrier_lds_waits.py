"""Every workgroup barrier of the kernels that reduce through LDS is reached with no LDS operation of
the wave still outstanding, on every control-flow path (CPU test: hipcc emits the gfx950 assembly).

__syncthreads() lets the compiler leave the LDS-count wait out where it takes LDS operations as
ordered across waves.  In the INT8 rho1 kernel a wave whose rows of the next row tile were all past
n_compute took the path without the wait, reached the barrier with its LDS write of the previous
tile's partial sums still outstanding, and wave 0 then read that slot stale now and then (a few points
of a pass wrong on the 60 000-batch grid).  The kernels now use an explicit wait before such
barriers; this test reads the assembly and checks, by a data-flow pass over each kernel's basic
blocks, that no s_barrier can be reached with an LDS read or write issued and not yet waited for
(lgkmcnt(0))."""
import os
import re
import shutil
import subprocess
import tempfile
from concurrent.futures import ThreadPoolExecutor

import pytest

from fast_rpa_eval_amd import build

# LDS memory operations (ds_bpermute / ds_swizzle move registers and are not counted)
LDS_MEM = re.compile(r"^ds_(read|write|load|store|add|sub|rsub|inc|dec|min|max|and|or|xor|mskor|cmpst|cmpswap|wrxchg|wrap)")
WAIT_LGKM0 = re.compile(r"^s_waitcnt\b(.*lgkmcnt\(0\)|\s+0\s*$)")
BRANCH = re.compile(r"^s_(c?branch\w*)\s+(\.LBB\w+)")


def barrier_report(asm):
    """[(kernel symbol, barriers, barriers reachable with an LDS operation outstanding)] for every
    kernel in `asm` that has a barrier."""
    out = []
    lines = asm.split("\n")
    i = 0
    while i < len(lines):
        m = re.match(r"^(_Z\S+):", lines[i])
        if not m:
            i += 1
            continue
        j = i + 1
        while j < len(lines) and not lines[j].startswith(".Lfunc_end"):
            j += 1
        blocks, labels = [[]], {}
        for raw in lines[i + 1:j]:
            lm = re.match(r"^(\.LBB\w+):", raw)
            if lm or raw.startswith("; %bb."):
                blocks.append([])
                if lm:
                    labels[lm.group(1)] = len(blocks) - 1
                continue
            s = raw.split(";")[0].strip()
            if s:
                blocks[-1].append(s)
        n = len(blocks)
        succ = []
        for b, ins in enumerate(blocks):
            nxt, fall = [], True
            for s in ins:
                bm = BRANCH.match(s)
                if bm:
                    nxt.append(labels[bm.group(2)])
                    fall = fall and bm.group(1) != "branch"
                if s.startswith(("s_endpgm", "s_setpc")):
                    fall = False
            if fall and b + 1 < n:
                nxt.append(b + 1)
            succ.append(nxt)

        def run(b, pending, flagged=None):
            for s in blocks[b]:
                if LDS_MEM.match(s):
                    pending = True
                elif WAIT_LGKM0.match(s):
                    pending = False
                elif s.startswith("s_barrier") and flagged is not None:
                    flagged.append(pending)
            return pending

        pend_in = [False] * n
        work = list(range(n))
        while work:
            b = work.pop()
            if run(b, pend_in[b]):
                for t in succ[b]:
                    if not pend_in[t]:
                        pend_in[t] = True
                        work.append(t)
        flags = []
        for b in range(n):
            run(b, pend_in[b], flags)
        if flags:
            out.append((m.group(1), len(flags), sum(flags)))
        i = j
    return out


def test_analysis_flags_a_path_without_the_wait():
    body = """_Zkernel:
	ds_write_b64 v1, v[2:3]
	s_cbranch_execz .LBB0_2
; %bb.1:
	s_waitcnt lgkmcnt(0)
	global_load_dword v4, v[5:6], off
.LBB0_2:
	s_barrier
	ds_read_b64 v[2:3], v1
	s_waitcnt lgkmcnt(0)
	s_barrier
	s_endpgm
.Lfunc_end0:
"""
    assert barrier_report(body) == [("_Zkernel", 2, 1)]  # the first barrier, via the skipping branch
    fixed = body.replace("\ts_cbranch_execz .LBB0_2\n", "\ts_waitcnt lgkmcnt(0)\n\ts_cbranch_execz .LBB0_2\n")
    assert barrier_report(fixed) == [("_Zkernel", 2, 0)]
    loop = """_Zloop:
.LBB1_1:
	s_barrier
	ds_write_b32 v1, v2
	s_cbranch_scc1 .LBB1_1
	s_endpgm
.Lfunc_end1:
"""
    assert barrier_report(loop) == [("_Zloop", 1, 1)]  # outstanding across the back edge


# the kernels whose workgroups combine partial results through LDS
SOURCES = ["rho1_ozaki.hip", "rho1_adjoint_ozaki.hip", "rho1_fp64.hip", "rho1_adjoint.hip", "rho1_bias2.hip",
           "rho1_lowp.hip"]


@pytest.mark.skipif(not os.path.exists(build.HIPCC), reason="hipcc not installed")
def test_library_barriers_wait_for_lds():
    tmp = tempfile.mkdtemp(prefix="frpa_asm_")
    try:
        def emit(src):
            out = os.path.join(tmp, src + ".s")
            r = subprocess.run([build.HIPCC, *build.CXXFLAGS, "--cuda-device-only", "-S", "-o", out,
                                os.path.join(build.CSRC, src)], capture_output=True, text=True)
            assert r.returncode == 0, r.stderr[-2000:]
            return src, barrier_report(open(out).read())
        with ThreadPoolExecutor(len(SOURCES)) as ex:
            reports = list(ex.map(emit, SOURCES))
    finally:
        shutil.rmtree(tmp, ignore_errors=True)
    bad = [(src, k, nb, nf) for src, rep in reports for k, nb, nf in rep if nf]
    n_kernels = sum(len(rep) for _, rep in reports)
    n_barriers = sum(nb for _, rep in reports for _, nb, _ in rep)
    print(f"{n_kernels} kernels, {n_barriers} barriers checked", flush=True)
    assert n_kernels >= 20 and n_barriers >= 100
    assert not bad, f"barriers reachable with LDS operations outstanding (source, kernel, barriers, flagged): {bad}"
