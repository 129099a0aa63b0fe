"""Collect the byte sizes of the reference's own test dumps, with the scalars its logs echo for them,
into a JSON fixture (dump_sizes.json next to this script).

The reference's dumps of FHI-aims runs (fragment_save_bin/case_<n>/first_order_rho.meta.test.nproc_4.bin
and sumup.test.nproc_4.bin) are Git-LFS pointer files: the data is not in the repository, but each
pointer records the file's size ("size <bytes>").  The direct-test runs print the scalars read from
those files (e.g. task2_matrix_shape_reorder/version2_gradient_rho_only/test_rho.log: "n_full_points =
355558" ... followed by "Run ../../fragment_save_bin/case_18 ...").  Together they pin the layout of
the format: the scalars fix every array's length, so the writer's byte count must equal the recorded
size.  This script reads those files as text (nothing is compiled or run) and writes

    {"cases": {"18": {"rho_scalars": {...}, "sumup_scalars": {...}, "rho_bytes": N, "sumup_bytes": M}, ...}}

Usage: python tests/golden/make_dump_sizes.py [/root/reference]
"""
import json
import os
import re
import sys

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
LOG = os.path.join(REF, "task2_matrix_shape_reorder", "version2_gradient_rho_only", "test_rho.log")
RHO_KEYS = ("n_my_batches_work", "n_full_points", "n_basis", "n_atoms", "n_max_compute_ham", "n_centers_basis_I",
            "n_max_batch_size")
SUMUP_KEYS = ("j_atom_begin", "j_atom_end", "n_full_points", "l_max_analytic_multipole", "index_cc_dim_0",
              "n_valid_points", "l_pot_max", "n_max_radial", "n_hartree_grid", "n_centers",
              "n_centers_hartree_potential", "n_atoms", "hartree_force_l_add", "n_species")


def log_blocks(path):
    """case -> the scalar lines printed before its "Run .../case_<n>" line, in order (the sum-up
    dump's scalars first, then the rho dump's, as the runner reads them)."""
    out, cur = {}, []
    for ln in open(path):
        m = re.match(r"\s*(\w+) = (-?\d+)\s*$", ln)
        if m:
            if m.group(1) == "j_atom_begin":
                cur = []
            cur.append((m.group(1), int(m.group(2))))
            continue
        m = re.search(r"fragment_save_bin/case_(\d+)", ln)
        if m and cur:
            out[m.group(1)] = cur
            cur = []
    return out


def main():
    cases = {}
    for case, lines in sorted(log_blocks(LOG).items(), key=lambda kv: int(kv[0])):
        keys = [k for k, _ in lines]
        vals = [v for _, v in lines]
        n_s = len(SUMUP_KEYS)
        assert tuple(keys[:n_s]) == SUMUP_KEYS and tuple(keys[n_s:n_s + len(RHO_KEYS)]) == RHO_KEYS, (case, keys)
        entry = {"sumup_scalars": dict(zip(SUMUP_KEYS, vals[:n_s])),
                 "rho_scalars": dict(zip(RHO_KEYS, vals[n_s:n_s + len(RHO_KEYS)]))}
        d = os.path.join(REF, "fragment_save_bin", f"case_{case}")
        for kind, name in (("rho", "first_order_rho.meta.test.nproc_4.bin"), ("sumup", "sumup.test.nproc_4.bin")):
            m = re.search(r"^size (\d+)$", open(os.path.join(d, name)).read(), re.M)
            entry[f"{kind}_bytes"] = int(m.group(1))
        cases[case] = entry
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "dump_sizes.json")
    with open(out, "w") as f:
        json.dump({"source": "fragment_save_bin/case_*/{first_order_rho.meta,sumup}.test.nproc_4.bin (LFS sizes); "
                             "task2_matrix_shape_reorder/version2_gradient_rho_only/test_rho.log (scalars)",
                   "cases": cases}, f, indent=1)
    print(f"{len(cases)} cases -> {out}")


if __name__ == "__main__":
    main()
