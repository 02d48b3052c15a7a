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


# ---- XML restart files (WriteConfig's "xml" and "base64" formats) ----------

def lj_restart_chunk(eps, sig, rcut):
    """The LJCut fix's restart text (FixPair::restartChunkPairParams,
    src/Fixes/FixPair.cpp:243-264): one ``<parameter handle="...">`` block per
    parameter matrix, in the handle order of its std::map ("eps", "rCut",
    "sig"), each the flattened ntypes x ntypes matrix, one value per line as
    std::ostream prints a double."""
    out = []
    for handle, mat in (("eps", eps), ("rCut", rcut), ("sig", sig)):
        out.append('<parameter handle="%s">\n' % handle)
        out.extend(_g(v) + "\n" for v in np.asarray(mat, np.float64).ravel())
        out.append("</parameter>\n")
    return "".join(out)


def rigid_restart_chunk(atoms4, style="TIP4P/2005", four_site=True):
    """FixRigid's restart text (FixRigid::restartChunk,
    src/Fixes/FixRigid.cu:1438-1456): the molecules' site ids (O, H, H, M;
    M = -1 for three-site water), the style and the model flag."""
    atoms4 = np.asarray(atoms4, np.int64).reshape(-1, 4)
    out = ['<atomsInMolecule n="%d">\n' % len(atoms4)]
    out.extend("%d %d %d %d\n" % tuple(r) for r in atoms4)
    out.append("</atomsInMolecule>\n")
    out.append("<style type='%s'>\n</style>\n" % style)
    out.append("<model fourSite='%s'/>\n" % ("true" if four_site else "false"))
    return "".join(out)


def _b64(a, dtype):
    import base64
    return base64.b64encode(np.ascontiguousarray(a, dtype=dtype).tobytes()).decode("ascii")


def xml_configuration(cfg, base64=False):
    """One ``<configuration>`` element in the layout of WriteConfig's
    writeXMLfile (text, src/WriteConfig.cpp:256-345) or writeXMLfileBase64
    (src/WriteConfig.cpp:91-160).

    cfg keys: turn, x, v, f ((n, 3) by atom), type, id, q, group_tag (n,),
    lo, hi (3,), periodic (3 bools), rcut, padding, dt, handles, masses (per
    type), fixes (list of (type, handle, restart text); written in the text
    format only, as the reference does), groups (dict handle -> bit; written
    in handle order like the reference's std::map), molecules (list of id
    lists) and, optionally, dimension (3)."""
    n = len(cfg["x"])
    per = "".join("1" if p else "0" for p in cfg["periodic"])
    out = ['<configuration turn="%d" numAtoms="%d" dimension="%d" periodic="%s" rCut="%f" padding="%f" dt="%f">\n'
           % (cfg["turn"], n, cfg.get("dimension", 3), per, cfg["rcut"], cfg["padding"], cfg["dt"])]
    lo, hi = np.asarray(cfg["lo"], np.float64), np.asarray(cfg["hi"], np.float64)
    if base64:
        b = [_b64(v, "<f8") for v in (lo[0], lo[1], lo[2], hi[0], hi[1], hi[2])]
        out.append('<bounds base64="1" xlo="%s" ylo="%s" zlo="%s" xhi="%s" yhi="%s" zhi="%s" />\n' % tuple(b))
    else:
        out.append('<bounds base64="0" xlo="%f" ylo="%f" zlo="%f" xhi="%f" yhi="%f" zhi="%f" />\n'
                   % (lo[0], lo[1], lo[2], hi[0], hi[1], hi[2]))
    # writeAtomParams (src/WriteConfig.cpp:36-53)
    out.append('<atomParams numTypes="%d">\n<handle>\n' % len(cfg["handles"]))
    out.extend(h + "\n" for h in cfg["handles"])
    out.append("</handle>\n<mass>\n")
    out.extend(_g(m) + "\n" for m in cfg["masses"])
    out.append("</mass>\n</atomParams>\n")
    if not base64:
        out.append("<fixes>\n")
        for ftype, handle, chunk in cfg.get("fixes", ()):
            out.append('<fix type="%s" handle="%s">\n' % (ftype, handle))
            out.append(chunk)
            out.append("</fix>\n")
        out.append("</fixes>\n")
    cols = (("position", cfg["x"], "<f8", "%f %f %f\n"), ("velocity", cfg["v"], "<f8", "%f %f %f\n"),
            ("force", cfg["f"], "<f8", "%f %f %f\n"), ("groupTag", cfg["group_tag"], "<u4", "%u\n"),
            ("type", cfg["type"], "<i4", "%d\n"), ("id", cfg["id"], "<i4", "%d\n"), ("q", cfg["q"], "<f8", "%f\n"))
    for tag, a, dt, fmt in cols:
        a = np.asarray(a)
        if base64:
            out.append('<%s base64="1">\n%s</%s>\n' % (tag, _b64(a, dt), tag))
        else:
            out.append("<%s>\n" % tag)
            if a.ndim == 2:
                out.extend(fmt % tuple(r) for r in a.astype(np.float64))
            else:
                out.extend(fmt % v for v in (a.astype(np.float64) if dt == "<f8" else a.astype(np.int64)))
            out.append("</%s>\n" % tag)
    # outputGroups / outputMolecules (src/WriteConfig.cpp:55-89)
    groups = cfg.get("groups", {"all": 1})
    out.append("<groupInfo>\n<groupHandles>\n")
    out.extend(h + "\n" for h in sorted(groups))
    out.append("</groupHandles>\n<groupBits>\n")
    out.extend("%d\n" % groups[h] for h in sorted(groups))
    out.append("</groupBits>\n</groupInfo>\n<molecules>\n")
    for ids in cfg.get("molecules", ()):
        out.append("<m>\n" + "".join("%d " % i for i in ids) + "</m>\n")
    out.append("</molecules>\n</configuration>\n")
    return "".join(out)


def write_xml(path, cfgs, base64=False, one_file_per_write=True):
    """A restart file of one or more configurations.  one_file_per_write
    follows a file name with ``*`` (``<data>`` and ``</data>`` each with a
    newline, writeXMLfile); otherwise the appended-trajectory layout of one
    file (``<data>`` written when the writer is made, ``</data>`` without a
    newline by WriteConfig::finish, src/WriteConfig.cpp:404-423)."""
    if isinstance(cfgs, dict):
        cfgs = [cfgs]
    body = "".join(xml_configuration(c, base64) for c in cfgs)
    with open(path, "w") as fh:
        fh.write("<data>\n" + body + ("</data>\n" if one_file_per_write else "</data>"))


def read_xml(path):
    """Configurations of a restart file written in either format: list of
    dicts with the keys xml_configuration takes (fix restart text is returned
    per fix as (type, handle, inner XML elements))."""
    import base64 as b64
    import xml.etree.ElementTree as ET

    root = ET.parse(path).getroot()
    out = []
    for c in root.iter("configuration"):
        n = int(c.get("numAtoms"))
        cfg = dict(turn=int(c.get("turn")), dimension=int(c.get("dimension")),
                   periodic=[ch == "1" for ch in c.get("periodic")], rcut=float(c.get("rCut")),
                   padding=float(c.get("padding")), dt=float(c.get("dt")))
        bnd = c.find("bounds")
        enc = bnd.get("base64") == "1"
        val = (lambda s: float(np.frombuffer(b64.b64decode(s), "<f8")[0])) if enc else float
        cfg["lo"] = np.array([val(bnd.get(k)) for k in ("xlo", "ylo", "zlo")])
        cfg["hi"] = np.array([val(bnd.get(k)) for k in ("xhi", "yhi", "zhi")])
        ap = c.find("atomParams")
        cfg["handles"] = ap.find("handle").text.split()
        cfg["masses"] = np.array([float(t) for t in ap.find("mass").text.split()])
        fx = c.find("fixes")
        cfg["fixes"] = [] if fx is None else [(f.get("type"), f.get("handle"), list(f)) for f in fx]
        for tag, key, dt, k in (("position", "x", "<f8", 3), ("velocity", "v", "<f8", 3), ("force", "f", "<f8", 3),
                                ("groupTag", "group_tag", "<u4", 1), ("type", "type", "<i4", 1),
                                ("id", "id", "<i4", 1), ("q", "q", "<f8", 1)):
            e = c.find(tag)
            if e.get("base64") == "1":
                a = np.frombuffer(b64.b64decode(e.text.strip()), dt).copy()
            else:
                a = np.array(e.text.split(), dtype=np.float64 if dt == "<f8" else np.int64).astype(dt)
            cfg[key] = a.reshape(n, 3) if k == 3 else a
        gi = c.find("groupInfo")
        cfg["groups"] = dict(zip(gi.find("groupHandles").text.split(),
                                 (int(t) for t in gi.find("groupBits").text.split())))
        cfg["molecules"] = [[int(t) for t in (m.text or "").split()] for m in c.find("molecules").iter("m")]
        out.append(cfg)
    return out
