"""Writes tests/golden/euler_C1_step5.npy: the view angles of the
reference's Euler_file/euler_C1_step5.lst (1649 views, a ~5 degree C1 sphere),
the list the full-size workload names.  Rows are (alt, az, phi) as float32,
the precision readEulerData parses them at (EMReader/DataReader.cpp:267-284,
sscanf "%f %f"; the third column is in the file but unused there).  Run in a
container that holds the reference: python tests/golden/make_euler_fixture.py
"""
import os

import numpy as np

SRC = "/root/reference/Euler_file/euler_C1_step5.lst"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "euler_C1_step5.npy")


def main() -> None:
    rows = []
    with open(SRC) as f:
        for line in f:
            if line.startswith("#") or not line.strip():
                continue
            rows.append([float(v) for v in line.split()[:3]])
    a = np.asarray(rows, dtype=np.float32)
    np.save(OUT, a)
    print(f"{OUT}: {a.shape[0]} views, alt {a[:, 0].min()}..{a[:, 0].max()}, "
          f"az {a[:, 1].min()}..{a[:, 1].max()}")


if __name__ == "__main__":
    main()
