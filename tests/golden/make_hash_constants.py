"""Regenerates tests/golden/hash_constants.txt: the 64 mixing constants of the
reference's PCG hash, read from the table in
/root/reference/nectar_render/csrc/random/include/hash.h (lines 17-34).  The
fixture is data (one hexadecimal word per line, table order)."""
import os
import re

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = "/root/reference/nectar_render/csrc/random/include/hash.h"

text = open(SRC).read()
body = re.search(r"HASH_CONSTANTS\[64\]\s*=\s*\{(.*?)\};", text, re.S).group(1)
words = [int(w, 16) for w in re.findall(r"0x([0-9A-Fa-f]+)u", body)]
assert len(words) == 64, len(words)
with open(os.path.join(HERE, "hash_constants.txt"), "w") as f:
    f.write("".join(f"0x{w:08X}\n" for w in words))
print(len(words), "constants")
