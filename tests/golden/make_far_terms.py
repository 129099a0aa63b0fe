"""Extract the far-field Cartesian term table held by the reference into a JSON fixture.

The reference ships its l <= 6 far-field dot product as an ahead-of-time unrolled sum
(fast_rap_src_dfpt_gpu_eval/sum_up_far_dist_potential_dot.hpp:12-210): one line per term,

    acc += p<i>(x) * p<j>(y) * p<k>(z) * Fp[<nn>] * multipole_c[<n>];

with ``if constexpr (l_max == L) return acc;`` after the last term of each l.  Those
(i, j, k, nn) exponents and the per-l term counts are the reference's own fixture for the
``index_cc`` columns 3-6 and ``n_cc_lm_ijk``.  This script reads that file as text (nothing is
compiled or run) and writes ``far_terms_l6.json`` next to itself:

    {"terms": [[i, j, k, nn], ...],  # 176 rows, in the reference's order n = 0..175
     "n_cc_lm_ijk": [1, 4, 12, 28, 57, 104, 176],
     "index_cc_dim_0_l7": 280}       # echoed by every far-field log (l_max_analytic_multipole = 7)

Usage: python tests/golden/make_far_terms.py [/root/reference]
"""
import json
import os
import re
import sys

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
SRC = os.path.join(REF, "fast_rap_src_dfpt_gpu_eval", "sum_up_far_dist_potential_dot.hpp")
LOG = os.path.join(REF, "task3_jit", "version5_sumup_far_part_with_jit_or_aot", "test_sum_up_far_part.log")

TERM = re.compile(r"acc \+= p(\d)\(x\) \* p(\d)\(y\) \* p(\d)\(z\) \* Fp\[(\d+)\] \* multipole_c\[(\d+)\];")
STOP = re.compile(r"if constexpr \(l_max == (\d+)\)")


def main():
    terms, n_cc = [], []
    with open(SRC) as f:
        for line in f:
            m = TERM.search(line)
            if m:
                i, j, k, nn, n = map(int, m.groups())
                assert n == len(terms), (n, len(terms))
                terms.append([i, j, k, nn])
                continue
            s = STOP.search(line)
            if s:
                assert int(s.group(1)) == len(n_cc)
                n_cc.append(len(terms))
    # The log echoes n_cc_lm_ijk[0:l_pot_max+1] and index_cc_dim_0 (task3 v5 log:12,33-39).
    with open(LOG) as f:
        log = f.read()
    logged = [int(v) for v in re.search(r"n_cc_lm_ijk\[0:l_pot_max\+1\]\n((?:\d+\n)+)", log).group(1).split()]
    assert logged == n_cc, (logged, n_cc)
    dim0 = int(re.search(r"index_cc_dim_0 = (\d+)", log).group(1))
    out = {"source": "fast_rap_src_dfpt_gpu_eval/sum_up_far_dist_potential_dot.hpp:12-210; "
                     "task3_jit/version5_sumup_far_part_with_jit_or_aot/test_sum_up_far_part.log:12,33-39",
           "terms": terms, "n_cc_lm_ijk": n_cc, "index_cc_dim_0_l7": dim0}
    dst = os.path.join(os.path.dirname(os.path.abspath(__file__)), "far_terms_l6.json")
    with open(dst, "w") as f:
        json.dump(out, f, separators=(",", ":"))
    print(f"wrote {dst}: {len(terms)} terms, n_cc_lm_ijk={n_cc}, index_cc_dim_0(l=7)={dim0}")


if __name__ == "__main__":
    main()
