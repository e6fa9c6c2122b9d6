"""Writes sobol_joe_kuo_table.h: primitive polynomials and initial direction
values (Joe & Kuo, new-joe-kuo-6.21201, as shipped with scipy) for the first
256 Sobol dimensions.  The engine derives the 32-bit direction numbers from
these at start-up (mp_api.hip, sobol_directions)."""
import os

import numpy as np
import scipy.stats as st

DIMS = 256


def main() -> None:
    path = os.path.join(os.path.dirname(st.__file__), "_sobol_direction_numbers.npz")
    with np.load(path) as d:
        poly = d["poly"][:DIMS].astype(np.uint64)
        vinit = d["vinit"][:DIMS].astype(np.uint64)
    degree = [int(p).bit_length() - 1 for p in poly]
    cols = max(max(degree), 1)
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "sobol_joe_kuo_table.h")
    with open(out, "w") as f:
        f.write("// Generated by gen_sobol_table.py from scipy's new-joe-kuo-6.21201 data.\n")
        f.write("#pragma once\n#include <stdint.h>\n")
        f.write(f"#define MP_SOBOL_TABLE_DIMS {DIMS}\n#define MP_SOBOL_TABLE_COLS {cols}\n")
        f.write("static const uint32_t mp_sobol_poly[MP_SOBOL_TABLE_DIMS] = {\n")
        for i in range(0, DIMS, 12):
            f.write("  " + ", ".join(str(int(x)) for x in poly[i:i + 12]) + ",\n")
        f.write("};\n")
        f.write("static const uint32_t mp_sobol_vinit[MP_SOBOL_TABLE_DIMS][MP_SOBOL_TABLE_COLS] = {\n")
        for i in range(DIMS):
            f.write("  {" + ", ".join(str(int(x)) for x in vinit[i, :cols]) + "},\n")
        f.write("};\n")


if __name__ == "__main__":
    main()
