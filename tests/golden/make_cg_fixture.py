"""Write tests/golden/cg_table_subset.json: the Clebsch-Gordan values held by the reference's
data file tf_pwa/cg_table.json for j1, j2 <= 2 (flattened to rows
[j1, j2, m1, m2, J, M, value]). Reads the reference tree; run once in the build container."""

import json
import os
import sys

SRC = "/root/reference/tf_pwa/cg_table.json"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cg_table_subset.json")


def main():
    with open(SRC) as f:
        t = json.load(f)
    rows = []
    for j1, a in t.items():
        for j2, b in a.items():
            if int(j1) > 2 or int(j2) > 2:
                continue
            for m1, c in b.items():
                for m2, d in c.items():
                    for J, e in d.items():
                        for M, v in e.items():
                            rows.append([int(j1), int(j2), int(m1), int(m2), int(J), int(M), v])
    with open(OUT, "w") as f:
        json.dump({"source": "tf_pwa/cg_table.json (j1, j2 <= 2)", "rows": rows}, f)
    print(len(rows), "rows ->", OUT)


if __name__ == "__main__":
    sys.exit(main())
