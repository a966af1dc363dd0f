"""Writes tests/golden/high_mat_sorted16.npz from the reference's shipped
METIS-sorted mesh (Assets/sorted_mesh/high_mat_sorted.16.msh, gmsh 2.2) and
its part file (high_mat_sorted.16.part, one part id per line).  Data only:
rest positions (float32 is enough for the connectivity checks; float64 kept
for the barrier-free Newton step), tets (0-based, int32) and part ids.

Run in the container that has the reference mounted:
    python tests/golden/make_high_mat_fixture.py /root/reference
"""
import os
import sys

import numpy as np


def read_msh(path):
    with open(path) as f:
        lines = f.read().split("\n")
    i = lines.index("$Nodes")
    nn = int(lines[i + 1])
    xs = np.array([list(map(float, l.split()[1:4])) for l in lines[i + 2:i + 2 + nn]])
    j = lines.index("$Elements")
    ne = int(lines[j + 1])
    tets = []
    for l in lines[j + 2:j + 2 + ne]:
        p = l.split()
        if p[1] == "4":  # 4-node tetrahedron
            ntags = int(p[2])
            tets.append([int(v) - 1 for v in p[3 + ntags:7 + ntags]])
    return xs, np.asarray(tets, np.int32)


def main(ref):
    d = os.path.join(ref, "Assets", "sorted_mesh")
    xs, tets = read_msh(os.path.join(d, "high_mat_sorted.16.msh"))
    part = np.loadtxt(os.path.join(d, "high_mat_sorted.16.part"), dtype=np.int32)
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "high_mat_sorted16.npz")
    np.savez_compressed(out, x_rest=xs, tets=tets, part=part)
    print("wrote", out, xs.shape, tets.shape, part.shape)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
