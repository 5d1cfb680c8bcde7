"""Writes tests/golden/sobol_tables.npz from the reference's Sobol' tables
(core/sobolmatrices.rs: SOBOL_MATRICES_32, VD_C_SOBOL_MATRICES,
VD_C_SOBOL_MATRICES_INV), read as text:

    python tests/golden/make_sobol_golden.py [path to rs_illume]

The fixture holds a SHA-256 digest of each full table (little-endian words in
table order) and a sample of entries (dimensions 0-3 and every 61st dimension
of the matrices; the first rows of the interval-to-index tables), so a test can
check tables built elsewhere against the reference's without the reference.
"""
import hashlib
import os
import re
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def main(ref_root):
    text = open(os.path.join(ref_root, "src", "core", "sobolmatrices.rs")).read()
    a = text.index("SOBOL_MATRICES_32")
    b = text.index("];", a)
    c = np.array([int(x, 16) for x in re.findall(r"0x([0-9a-fA-F]+)", text[a:b])], np.uint32)
    assert c.size == 1024 * 52

    def rows(prefix, count):
        out = []
        for m in range(1, count + 1):
            i = text.index("const %s%d:" % (prefix, m))
            j = text.index("];", i)
            body = text[text.index("=", i):j]
            out.append(np.array([int(x, 16) for x in re.findall(r"0x([0-9a-fA-F]+)_u64", body)], np.uint64))
        return out

    vdc = rows("M", 25)
    inv = rows("MI", 26)

    def digest(arrs):
        h = hashlib.sha256()
        for x in arrs:
            h.update(np.ascontiguousarray(x).astype(x.dtype.newbyteorder("<")).tobytes())
        return np.frombuffer(h.digest(), np.uint8)

    c2 = c.reshape(1024, 52)
    dims = np.array(sorted(set(range(4)) | set(range(0, 1024, 61))), np.int32)
    np.savez(os.path.join(HERE, "sobol_tables.npz"),
             matrices_sha256=digest([c]), vdc_sha256=digest(vdc), vdc_inv_sha256=digest(inv),
             sample_dims=dims, sample_matrices=c2[dims],
             vdc_m1_m4=np.concatenate(vdc[:4]), vdc_inv_m1_m4=np.concatenate(inv[:4]),
             vdc_lengths=np.array([len(x) for x in vdc], np.int32),
             vdc_inv_lengths=np.array([len(x) for x in inv], np.int32))
    print("wrote", os.path.join(HERE, "sobol_tables.npz"))


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
