"""Writes python_bindings.txt: every clibrebound.reb_* function the reference's
Python package binds, with the module that binds it, read as text from the
package's sources (nothing is imported or run).

    python tests/golden/list_python_bindings.py [/root/reference/rebound]
"""
import os
import re
import sys

src = sys.argv[1] if len(sys.argv) > 1 else "/root/reference/rebound"
pat = re.compile(r"clibrebound\.(reb_\w+)")
rows = set()
for name in sorted(os.listdir(src)):
    if name.endswith(".py"):
        with open(os.path.join(src, name)) as f:
            for fn in pat.findall(f.read()):
                rows.add((name, fn))
out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "python_bindings.txt")
with open(out, "w") as f:
    f.write("# module function: clibrebound.reb_* calls of the reference's Python package\n")
    for name, fn in sorted(rows):
        f.write("%s %s\n" % (name, fn))
print("wrote", out, len(rows))
