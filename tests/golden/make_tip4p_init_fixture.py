"""Writes tests/golden/tip4p_init_excerpt.out: the header, box and the first
and last four molecules of the reference's examples/tip4p_2005_speed/
tip4p_init_8000.out (data rows copied verbatim; natoms rewritten to 32), and
tip4p_init_8000.json with the full file's natoms, box diagonal and the first
and last site rows (nm) for the reader test."""
import json
import os
import re

SRC = "/root/reference/examples/tip4p_2005_speed/tip4p_init_8000.out"
HERE = os.path.dirname(os.path.abspath(__file__))

lines = open(SRC).read().splitlines()
hdr = next(i for i, s in enumerate(lines) if "natoms=" in s)
natoms = int(lines[hdr].split()[1])
xs = next(i for i, s in enumerate(lines) if s.strip().startswith("x ("))
rows = lines[xs + 1: xs + 1 + natoms]
row = re.compile(r"\{\s*([-+0-9.eE]+),\s*([-+0-9.eE]+),\s*([-+0-9.eE]+)\s*\}")
box = [float(row.search(lines[hdr + 2 + a]).group(a + 1)) for a in range(3)]


def vals(s):
    m = row.search(s)
    return [float(m.group(k)) for k in (1, 2, 3)]


keep = rows[:16] + rows[-16:]
with open(os.path.join(HERE, "tip4p_init_excerpt.out"), "w") as f:
    f.write("\n".join(lines[:hdr]) + "\n")
    f.write(lines[hdr].replace(f"natoms={natoms:10d}", f"natoms={32:10d}") + "\n")
    f.write("\n".join(lines[hdr + 1: xs]) + "\n")
    f.write(f"   x (32x3):\n")
    for k, s in enumerate(keep):
        f.write(re.sub(r"\[\s*\d+\]", f"[{k:5d}]", s) + "\n")
json.dump(dict(natoms=natoms, box_nm=box, first_nm=[vals(s) for s in rows[:4]],
               last_nm=[vals(s) for s in rows[-4:]]),
          open(os.path.join(HERE, "tip4p_init_8000.json"), "w"), indent=1)
print("wrote fixtures", natoms, box)
