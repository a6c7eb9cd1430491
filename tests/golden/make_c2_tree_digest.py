"""Regenerates tests/golden/c2_mesh_tree.txt: the SHA-256 digest of the oracle's
binned-SAH tree (child-pair node array + leaf order) of the 1,003,520-triangle
displaced icosphere, followed by its node and triangle counts.  The GPU test
checks the product's host and device builders against the same digest."""
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
sys.path.insert(0, os.path.dirname(HERE))

from oracle import pyoracle  # noqa: E402
import test_oracle_pins as pins  # noqa: E402

digest, n = pins.c2_tree_digest(pyoracle)
with open(os.path.join(HERE, "c2_mesh_tree.txt"), "w") as f:
    f.write(f"{digest} {n - 1} {n}\n")
print(digest, n)
