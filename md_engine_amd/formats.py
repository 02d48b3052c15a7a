"""Input formats for the water workloads.

``read_gmx_dump`` parses the GROMACS ``gmx dump`` text frames that the
reference's water examples start from (the ``tip4p_init_*.out`` files, read by
``read_gromacs_trajectory`` in examples/e3b3_speed/e3b3_8000.py:17-73): per
frame a header line with ``natoms=``, a ``box (3x3):`` block whose diagonal is
the orthorhombic box, then ``x (Nx3):`` and optionally ``v (Nx3):`` blocks of
``name[ i]={ a, b, c}`` rows.  Lengths are nm and velocities nm/ps in the
file; they are returned in A and A/fs (x10 and x0.01, as the example does).

``tip4p_init_box`` turns such a frame into the same dict layout as
``systems.tip4p2005_box`` (sites O, H1, H2, M per molecule, in file order).
With ``idealize`` each molecule is placed on the exact rigid TIP4P/2005
geometry (O kept, the HOH plane and bisector taken from the file, |OH|, |HH|
and the M site set from the constraint values), since the file rounds
coordinates to 6 significant digits and SETTLE needs an exactly rigid start.
"""
from __future__ import annotations

import re

import numpy as np

from . import systems

_ROW = re.compile(r"\{\s*([-+0-9.eE]+),\s*([-+0-9.eE]+),\s*([-+0-9.eE]+)\s*\}")


def read_gmx_dump(path, n_frames=1):
    """Frames of a ``gmx dump`` text file: list of dicts with natoms, box (3,)
    in A, x (natoms, 3) in A and, when present, v (natoms, 3) in A/fs."""
    frames = []
    with open(path) as f:
        lines = iter(f.read().splitlines())
    cur = None
    for line in lines:
        s = line.strip()
        if s.startswith("natoms="):
            if cur is not None:
                frames.append(cur)
                if len(frames) == n_frames:
                    return frames
            cur = dict(natoms=int(s.split()[1]))
        elif cur is None:
            continue
        elif s.startswith("box (3x3)"):
            rows = [_ROW.search(next(lines)) for _ in range(3)]
            cur["box"] = np.array([float(rows[a].group(a + 1)) for a in range(3)]) * 10.0
        elif s.startswith("x (") or s.startswith("v ("):
            key = s[0]
            n = cur["natoms"]
            arr = np.empty((n, 3))
            for i in range(n):
                m = _ROW.search(next(lines))
                if m is None:
                    raise ValueError(f"{path}: bad {key} row {i}")
                arr[i] = [float(m.group(1)), float(m.group(2)), float(m.group(3))]
            cur[key] = arr * (10.0 if key == "x" else 0.01)
    if cur is not None:
        frames.append(cur)
    return frames[:n_frames]


def idealize_tip4p2005(x, L):
    """Exact rigid TIP4P/2005 molecules from approximate sites (n, 4, 3):
    O kept; the bisector and the HOH plane follow the file's H positions."""
    p = systems.TIP4P2005
    O = x[:, 0]
    h1 = x[:, 1] - O
    h2 = x[:, 2] - O
    h1 -= L * np.rint(h1 / L)
    h2 -= L * np.rint(h2 / L)
    bis = h1 + h2
    bis /= np.linalg.norm(bis, axis=1)[:, None]
    perp = h1 - h2
    perp -= (perp * bis).sum(1)[:, None] * bis
    perp /= np.linalg.norm(perp, axis=1)[:, None]
    half = 0.5 * p["r_HH"]
    mid = np.sqrt(p["r_OH"] ** 2 - half ** 2)
    H1 = O + mid * bis + half * perp
    H2 = O + mid * bis - half * perp
    M = O + 0.5 * (1.0 - p["gamma"]) * ((H1 - O) + (H2 - O))
    return np.stack([O, H1, H2, M], axis=1)


def tip4p_init_box(path, idealize=True):
    """A TIP4P/2005 water box (systems.tip4p2005_box layout) from the first
    frame of a ``tip4p_init_*.out`` file."""
    fr = read_gmx_dump(path, 1)[0]
    n = fr["natoms"]
    if n % 4:
        raise ValueError(f"{path}: {n} sites is not a whole number of 4-site waters")
    nmol = n // 4
    L = fr["box"]
    x = fr["x"].reshape(nmol, 4, 3)
    if idealize:
        x = idealize_tip4p2005(x, L)
    x = x.reshape(-1, 3) % L
    p = systems.TIP4P2005
    q = np.tile([0.0, p["q_H"], p["q_H"], p["q_M"]], nmol)
    typ = np.tile([0, 1, 1, 2], nmol).astype(np.int32)
    mass = np.tile([p["m_O"], p["m_H"], p["m_H"], p["m_M"]], nmol)
    mol = np.arange(n, dtype=np.int32).reshape(nmol, 4)
    out = dict(x=x, q=q, type=typ, mass=mass, mol=mol, L=L.copy(), lo=np.zeros(3))
    if "v" in fr:
        out["v"] = fr["v"]
    return out


def _g(v):
    """a double as std::ostream prints it by default (6 significant digits)"""
    return "%g" % v


def write_xyz(path, x, atomic_nums, lo, hi, v=None, f=None, append=True):
    """Append one frame in the layout of WriteConfig's xyz writer
    (writeXYZFile, src/WriteConfig.cpp:206-254): the atom count, then
    ``bounds lo (x, y, z) hi (x, y, z)`` and one line per atom with its atomic
    number (or type) and position, then velocity and force when given; the
    frame ends with a newline.  x, v, f: (n, 3) by atom id."""
    x = np.asarray(x, np.float64)
    n = len(x)
    vec = lambda a: "(" + ", ".join(_g(c) for c in a) + ")"  # noqa: E731  Vector::asStr
    out = [str(n) + "\n" + "bounds lo " + vec(lo) + " hi " + vec(hi)]
    for i in range(n):
        row = "\n" + str(int(atomic_nums[i])) + " " + " ".join(_g(c) for c in x[i])
        if v is not None:
            row += " " + " ".join(_g(c) for c in v[i])
        if f is not None:
            row += " " + " ".join(_g(c) for c in f[i])
        out.append(row)
    out.append("\n")
    with open(path, "a" if append else "w") as fh:
        fh.write("".join(out))


def read_xyz(path):
    """frames written by write_xyz: list of (atomic numbers, (n, k) values, lo, hi)"""
    frames = []
    with open(path) as fh:
        lines = fh.read().split("\n")
    i = 0
    while i < len(lines) and lines[i].strip():
        n = int(lines[i])
        b = re.findall(r"\(([^)]*)\)", lines[i + 1])
        lo = np.array([float(t) for t in b[0].split(",")])
        hi = np.array([float(t) for t in b[1].split(",")])
        rows = [ln.split() for ln in lines[i + 2:i + 2 + n]]
        nums = np.array([int(r[0]) for r in rows])
        vals = np.array([[float(t) for t in r[1:]] for r in rows])
        frames.append((nums, vals, lo, hi))
        i += 2 + n
    return frames
