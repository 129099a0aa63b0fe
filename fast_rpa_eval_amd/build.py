"""Build libfrpa.so (HIP kernels + C-ABI) for gfx950, in-tree.

    python -m fast_rpa_eval_amd.build [-j N] [-v]

Steps: regenerate csrc/far_terms.inc (the compile-time Cartesian term table, from
multipole_tables.py), compile every csrc/*.hip with hipcc --offload-arch=gfx950 in parallel into
_build/, link fast_rpa_eval_amd/libfrpa.so.
"""
import argparse
import concurrent.futures as cf
import glob
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libfrpa.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
TERMS_LMAX = 10

CXXFLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    # packed FP32 instructions are never wanted in this library
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-munsafe-fp-atomics",  # hardware FP64 atomic add for the adjoint scatter
    "-pthread",
    "-Wall", "-Wno-unused-function", "-Wno-unused-command-line-argument",
    f"-I{os.path.join(ROOT, 'include')}", f"-I{CSRC}",
]


def far_terms_header():
    sys.path.insert(0, ROOT)
    from fast_rpa_eval_amd.multipole_tables import cartesian_terms
    rows = cartesian_terms(TERMS_LMAX)
    ncc = [sum(1 for r in rows if r[0] <= l) for l in range(TERMS_LMAX + 1)]
    lines = [
        "// GENERATED by fast_rpa_eval_amd/build.py from multipole_tables.cartesian_terms -- do not edit.",
        "// Cartesian far-field terms (l, m, i, j, k, nn), rows ordered l, m = -l..l, then (j, k).",
        "#pragma once",
        "namespace frpa {",
        "struct FarTerm { int l, m, i, j, k, nn; };",
        f"constexpr int FAR_TAB_LMAX = {TERMS_LMAX};",
        "constexpr int FAR_NCC[FAR_TAB_LMAX + 1] = {" + ", ".join(map(str, ncc)) + "};",
        f"constexpr FarTerm FAR_TERMS[{len(rows)}] = {{",
    ]
    for (l, m, i, j, k, nn, _) in rows:
        lines.append(f"    {{{l}, {m}, {i}, {j}, {k}, {nn}}},")
    lines += ["};", "} // namespace frpa", ""]
    return "\n".join(lines)


def write_if_changed(path, text):
    old = open(path).read() if os.path.exists(path) else None
    if old != text:
        with open(path, "w") as f:
            f.write(text)


def compile_one(src, verbose):
    obj = os.path.join(BUILD, os.path.basename(src) + ".o")
    cmd = [HIPCC, *CXXFLAGS, "-c", src, "-o", obj]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    out = "\n".join(l for l in (r.stdout + r.stderr).splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src}:\n{out}")
    if out.strip() and verbose:
        print(out)
    return obj


def build(jobs=None, verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    write_if_changed(os.path.join(CSRC, "far_terms.inc"), far_terms_header())
    srcs = sorted(glob.glob(os.path.join(CSRC, "*.hip")))
    jobs = jobs or min(8, os.cpu_count() or 1)
    with cf.ThreadPoolExecutor(jobs) as ex:
        objs = list(ex.map(lambda s: compile_one(s, verbose), srcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-pthread", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n" + r.stdout + r.stderr)
    os.replace(tmp, LIB)
    return LIB


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("-j", type=int, default=None)
    ap.add_argument("-v", action="store_true")
    a = ap.parse_args()
    print(build(a.j, a.v))


if __name__ == "__main__":
    main()
