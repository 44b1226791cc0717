"""Writes f03_bound_symbols.txt: the C symbol names that the reference's
Fortran module binds (every bind(C, name="...") in src/f03/), one per line.
Run in a checkout that has the reference at /root/reference; the output is
data (a list of names) that tests/test_cabi.py checks against the library's
exports."""
import os
import re

SRC = "/root/reference/src/f03"
names = set()
for fn in sorted(os.listdir(SRC)):
    if fn.endswith((".f03", ".f90", ".F90")):
        text = open(os.path.join(SRC, fn)).read()
        names.update(re.findall(r'bind\s*\(\s*c\s*,\s*name\s*=\s*"(\w+)"', text, flags=re.I))
out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "f03_bound_symbols.txt")
with open(out, "w") as f:
    f.write("\n".join(sorted(names)) + "\n")
print(len(names), "names ->", out)
