"""Writes csrc/sobol_init.h: the primitive polynomials and initial direction
numbers of the first 1024 Sobol' dimensions (Joe and Kuo, "Constructing
Sobol sequences with better two-dimensional projections", SIAM J. Sci.
Comput. 30, 2008; the new-joe-kuo-6.21201 set, read from the copy SciPy
ships). The library expands them into the generator matrices at run time;
dimension 0 is the van der Corput sequence.

    python -m rs_illume_amd.gen_sobol_init
"""
import os

import numpy as np

N_DIMS = 1024
MAX_DEG = 13


def load():
    import scipy
    z = np.load(os.path.join(os.path.dirname(scipy.__file__), "stats", "_sobol_direction_numbers.npz"))
    poly = z["poly"][:N_DIMS].astype(np.int64)
    vinit = z["vinit"][:N_DIMS].astype(np.int64)
    if int(vinit[:, MAX_DEG:].max(initial=0)) != 0:
        raise ValueError("direction numbers beyond degree %d" % MAX_DEG)
    return poly, vinit[:, :MAX_DEG]


def main():
    poly, vinit = load()
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "csrc", "sobol_init.h")
    lines = ["// Generated by rs_illume_amd/gen_sobol_init.py: Joe-Kuo (new-joe-kuo-6.21201)",
             "// primitive polynomials (bit s = x^s, ..., bit 0 = 1) and initial direction",
             "// numbers m_1..m_s of the first %d Sobol' dimensions (dimension 0: van der" % N_DIMS,
             "// Corput, polynomial 1).",
             "#pragma once",
             "#include <stdint.h>",
             "",
             "static const uint16_t kSobolPoly[%d] = {" % N_DIMS]
    for i in range(0, N_DIMS, 16):
        lines.append("    " + ", ".join(str(int(v)) for v in poly[i:i + 16]) + ",")
    lines.append("};")
    lines.append("static const uint16_t kSobolInit[%d][%d] = {" % (N_DIMS, MAX_DEG))
    for row in vinit:
        lines.append("    {" + ", ".join(str(int(v)) for v in row) + "},")
    lines.append("};")
    with open(out, "w") as f:
        f.write("\n".join(lines) + "\n")
    print("wrote", out)


if __name__ == "__main__":
    main()
