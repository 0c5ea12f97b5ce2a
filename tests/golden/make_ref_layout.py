#!/usr/bin/env python3
"""Writes tests/golden/ref_layout.json: sizeof/offsetof of the reference data
model (src/common/geauxdock.h, geauxdock_soa.h), measured by compiling a tiny
program against the reference headers with g++.  The drop-in header
include/geauxdock_dropin.h must reproduce every number (test_dropin_layout).
Run in a container where /root/reference is present:
    python tests/golden/make_ref_layout.py
"""
import json
import os
import subprocess
import tempfile

REF = "/root/reference/src/common"
HERE = os.path.dirname(os.path.abspath(__file__))
FIELDS = {
    "Ligand": ["x1", "x2", "x3", "t3", "c3", "kde_begin_idx", "kde_end_idx", "center", "lig_natom"],
    "Protein": ["t", "x", "y", "z", "cdc", "hpp", "ele", "c", "seq3r", "pocket_center", "prt_npoint"],
    "Psp": ["psp"], "Kde": ["t", "x", "kde_npoint"], "Mcs": ["x", "y", "z", "ncol", "tcc"],
    "Mcs_R": ["tcc"], "Mcs_ELL": ["ncol", "tcc"], "Mcs_CSR": ["tcc", "npoint"],
    "EnePara": ["p1a", "p2a", "lj0", "lj1", "el1", "el0", "a1", "b1", "pmf0", "pmf1", "hdb0", "hdb1",
                "hpp", "hpl0", "hpl1", "hpl2", "kde2", "kde3", "a_para", "b_para", "w"],
    "ComplexSize": ["n_lig", "n_rep", "mcs_nrow"], "Temp": ["t", "minus_beta", "order"],
    "ReplicaMC": ["idx_rep", "idx_tmp", "movematrix", "energy", "step", "is_accept"],
    "McPara": ["steps_total", "steps_per_dump", "steps_per_exchange", "move_scale", "outputdir", "outputfile"],
    "Record": ["replica", "next_entry"],
    "Complex": ["lig", "prt", "psp", "kde", "mcs", "mcs_r", "mcs_ell", "mcs_csr", "enepara", "temp",
                "replica", "size", "mcpara", "lig_file", "weight_file", "rep_begin", "rep_end", "signal"],
    "ParaT": ["rep_begin", "rep_end", "n_rep", "record_sz"],
}


def program(include):
    lines = ["#include <cstdio>", "#include <cstddef>", include, "int main(){"]
    for s, fs in FIELDS.items():
        lines.append('printf("%s sizeof %%zu\\n", sizeof(%s));' % (s, s))
        for f in fs:
            lines.append('printf("%s %s %%zu\\n", offsetof(%s, %s));' % (s, f, s, f))
    lines.append("}")
    return "\n".join(lines)


def measure(include, flags):
    with tempfile.TemporaryDirectory() as d:
        src = os.path.join(d, "l.cpp")
        exe = os.path.join(d, "l")
        open(src, "w").write(program(include))
        subprocess.run(["g++", "-w", "-std=c++17"] + flags + [src, "-o", exe], check=True)
        out = subprocess.run([exe], capture_output=True, text=True, check=True).stdout
    res = {}
    for line in out.splitlines():
        s, f, v = line.split()
        res["%s.%s" % (s, f)] = int(v)
    return res


if __name__ == "__main__":
    ref = measure("#include <geauxdock.h>", ["-I" + REF])
    json.dump(ref, open(os.path.join(HERE, "ref_layout.json"), "w"), indent=0, sort_keys=True)
    print("wrote", len(ref), "entries")
