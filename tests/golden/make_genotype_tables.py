"""Writes genotype_tables.npy: the four 256x4 decode tables of the reference
(src/genotype_tables.hpp: additive, heterozygous, dominant, recessive), as
int8 data, for the decode known-answer tests.  Reads the reference header as
text (run in the build container, where /root/reference exists)."""
import os
import re
import sys

import numpy as np

SRC = "/root/reference/src/genotype_tables.hpp"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "genotype_tables.npy")


def main():
    text = open(SRC).read()
    names = ["_ADDITIVE_GENOTYPE_TABLE", "_HETEROZYGOUS_GENOTYPE_TABLE",
             "_DOMINANT_GENOTYPE_TABLE", "_RECESSIVE_GENOTYPE_TABLE"]
    tables = []
    for name in names:
        start = text.index(name)
        body = text[text.index("{", start) + 1:text.index("};", start)]
        rows = re.findall(r"\{([^}]*)\}", body)
        vals = [[float(x) for x in r.split(",")] for r in rows]
        assert len(vals) == 256 and all(len(v) == 4 for v in vals), name
        tables.append(vals)
    arr = np.array(tables, dtype=np.int8)
    np.save(OUT, arr)
    print("wrote", OUT, arr.shape)


if __name__ == "__main__":
    sys.exit(main())
