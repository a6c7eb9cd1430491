"""Regenerates tests/golden/philox_rocrand.txt from rocRAND's own host-callable
Philox (rocrand_init + rocrand_uniform), the open-source generator whose
(seed, subsequence, offset) semantics the oracle restates.  Needs hipcc."""
import os
import subprocess
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))

with tempfile.TemporaryDirectory() as d:
    exe = os.path.join(d, "philox_ref")
    subprocess.run(["hipcc", "-O1", "-o", exe, os.path.join(HERE, "rocrand_philox_ref.cpp")], check=True)
    out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout
with open(os.path.join(HERE, "philox_rocrand.txt"), "w") as f:
    f.write(out)
print(len(out.splitlines()), "vectors")
