"""Writes the data fixtures that pin the host side against the reference's
own files (run in a container where the reference tree is mounted; the
outputs are committed, the tests never read the reference):

  abi_structs.json   field lists (type, name, array length) of the structs
                     of SOURCE/header_MDOODZ.h that cross the drop-in ABI
  flow_laws.json     table entries of SOURCE/FlowLaws.c used by the
                     synthetic models (power law 13, 23, 40; linear 40;
                     exponential 40), as published (SI units, unscaled)
  matrix_pins.npz    structure of the reference's exported matrices
                     MATLAB_visualisation/Jacobian.mat (Newton blocks of a
                     50 x 50 cell model) and DeCoupledMatrix.mat (Picard
                     blocks): shapes, per-row non-zero counts, and the full
                     non-zero pattern of the Jacobian's A block

Usage: python tests/golden/gen_reference_fixtures.py [/root/reference]
"""
import json
import os
import re
import sys

import numpy as np
import scipy.io

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
import header_decls  # noqa: E402

STRUCTS = ["scale", "mat_prop", "BC", "BCT", "SparseMat", "params", "Nparams", "grid"]


def abi(ref):
    text = open(os.path.join(ref, "SOURCE", "header_MDOODZ.h")).read()
    return {s: header_decls.struct_fields(text, s) for s in STRUCTS}


def _function_body(text, name):
    i = text.index("void " + name + "(")
    j = text.index("{", i)
    depth = 0
    for k in range(j, len(text)):
        if text[k] == "{":
            depth += 1
        elif text[k] == "}":
            depth -= 1
            if depth == 0:
                return text[j:k + 1]
    raise ValueError(name)


def _case(body, number):
    m = re.search(r"case\s+%d\s*:(.*?)break\s*;" % number, body, flags=re.S)
    if not m:
        raise KeyError(number)
    out = {}
    for fld, val in re.findall(r"mat->(\w+)\[k\]\s*=\s*([-+0-9.eE*]+)\s*;", m.group(1)):
        x = 1.0
        for factor in val.split("*"):  # numeric literals and their products
            x *= float(factor)
        out[fld] = x
    return out


def flow_laws(ref):
    text = open(os.path.join(ref, "SOURCE", "FlowLaws.c")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    text = re.sub(r"//[^\n]*", "", text)
    return {
        "pwl": {str(n): _case(_function_body(text, "ReadDataPowerLaw"), n) for n in (13, 23, 40)},
        "lin": {str(n): _case(_function_body(text, "ReadDataLinear"), n) for n in (40,)},
        "exp": {str(n): _case(_function_body(text, "ReadDataExponential"), n) for n in (40,)},
    }


def matrix_pins(ref):
    out = {}
    for tag, fname in (("jac", "Jacobian.mat"), ("pic", "DeCoupledMatrix.mat")):
        d = scipy.io.loadmat(os.path.join(ref, "MATLAB_visualisation", fname))
        for blk in ("MA", "MB", "MC", "MD"):
            M = d[blk].tocsr()
            M.sum_duplicates()
            out["%s_%s_shape" % (tag, blk)] = np.array(M.shape, np.int64)
            out["%s_%s_rownnz" % (tag, blk)] = np.diff(M.indptr).astype(np.int16)
            if tag == "jac" and blk == "MA":
                out["jac_MA_indptr"] = M.indptr.astype(np.int32)
                out["jac_MA_indices"] = M.indices.astype(np.int32)
    return out


def main():
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    with open(os.path.join(HERE, "abi_structs.json"), "w") as f:
        json.dump(abi(ref), f, indent=0)
    with open(os.path.join(HERE, "flow_laws.json"), "w") as f:
        json.dump(flow_laws(ref), f, indent=1, sort_keys=True)
    np.savez_compressed(os.path.join(HERE, "matrix_pins.npz"), **matrix_pins(ref))


if __name__ == "__main__":
    main()
