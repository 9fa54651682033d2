#!/usr/bin/env python3
"""Write ``oracle/oracle_bg_tables.h`` from the reference's C table text.

The oracle takes its base graphs from a different source than the product:
the product header is generated from the 3GPP .mat1 matrices, the oracle's
from the arrays in c_lib/ldpc_tables.c:30-68 (read as text, numbers only).  A
CPU test checks the two agree entry by entry, which pins both.

Run in the development container only (it reads /root/reference); the
generated header is committed.
"""
import os
import re
import sys

HERE = os.path.dirname(os.path.abspath(__file__))


def parse_arrays(text):
    arrays = {}
    # name[...]... = { numbers }  (2-D arrays are flattened in row order)
    for m in re.finditer(r"uint16_t\s+(\w+)\s*(\[[^=]*\])\s*=\s*\{(.*?)\};",
                         text, re.S):
        name, body = m.group(1), m.group(3)
        body = body.replace("BG1_MAX_NON_NULL_ELEMENTS", "316")
        body = body.replace("BG2_MAX_NON_NULL_ELEMENTS", "197")
        nums = [int(x) for x in re.findall(r"-?\d+", body)]
        arrays[name] = nums
    return arrays


def main():
    src = sys.argv[1] if len(sys.argv) > 1 else \
        "/root/reference/c_lib/ldpc_tables.c"
    with open(src) as f:
        arrays = parse_arrays(f.read())
    spec = {
        1: (46, 316),
        2: (42, 197),
    }
    out = ["/* Generated by oracle/gen_oracle_tables.py from the numeric arrays",
           " * of c_lib/ldpc_tables.c.  Oracle-only data; do not edit. */",
           "#ifndef ORACLE_BG_TABLES_H", "#define ORACLE_BG_TABLES_H",
           "#include <stdint.h>", ""]
    for bgn, (rows, nnz) in spec.items():
        off = arrays[f"BG{bgn}_HBG_offset_array"]
        col = arrays[f"BG{bgn}_HBG_array"]
        vij = arrays[f"BG{bgn}_HBG_Vij_array"]
        assert len(off) == rows + 1 and off[-1] == nnz, (bgn, len(off))
        assert len(col) == nnz, (bgn, len(col))
        assert len(vij) == 8 * nnz, (bgn, len(vij))
        out.append(f"static const uint16_t orc_bg{bgn}_row_off[{rows + 1}] = {{"
                   + ",".join(map(str, off)) + "};")
        out.append(f"static const uint16_t orc_bg{bgn}_col[{nnz}] = {{"
                   + ",".join(map(str, col)) + "};")
        out.append(f"static const uint16_t orc_bg{bgn}_vij[8 * {nnz}] = {{")
        for i in range(8):
            out.append("  " + ",".join(map(str, vij[i * nnz:(i + 1) * nnz])) + ",")
        out.append("};")
        out.append("")
    out.append("#endif")
    path = os.path.join(HERE, "oracle_bg_tables.h")
    with open(path, "w") as f:
        f.write("\n".join(out) + "\n")
    print("wrote", path)


if __name__ == "__main__":
    main()
