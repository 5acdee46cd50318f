"""NEMO-3D `nd_*` files: a text header followed by binary records, and the
ASCII form written by the reference's binary-to-ASCII converter.

Header (writeHeader, NEMO_3D/nemo3d/src/io/fileformat.c:182-205):

    <HDR>
    record format        = {real kx, real ky, real kz, real E}
    record dimension 1   = 6400       # number of eigenvalues
    user data 1          =# not used
    <\\HDR>

then prod(dimensions) records of the listed fields, native little-endian:
`real` 8 bytes, `float` 4, `int` 4 (parseRecordFormat, fileformat.c:53-100).
The ASCII form (convertBinToAscii, fileformat.c:875-990) repeats the header
as read back (the record format and user data lines gain the space that
followed '='), then one line per record: reals as %21.12e, floats as
%16.7e, ints as %10d.

Writers for the files the solver path produces: eigenvalues (nd_Ek),
eigenvectors (nd_evec_k), atom positions (nd_rAtom), neighbour indices
(nd_nbrIndx), atom types (nd_aType) and displacements (nd_disp)
(i_3dout_files.c:323, 1342, 1889, 2009).
"""
import re

import numpy as np

_TYPES = {"real": "<f8", "float": "<f4", "int": "<i4"}


class NdFile:
    def __init__(self, fmt, dims, user, data):
        self.fmt = fmt          # record format text, e.g. "{real x, real y}"
        self.dims = dims        # [(count, label)]
        self.user = user        # [text after '=']
        self.data = data        # structured numpy array, one row per record

    def field(self, name):
        return self.data[name]


def parse_format(fmt):
    """[(name, dtype)] of a single-group record format."""
    m = re.fullmatch(r"\s*\{([^}]*)\}\s*", fmt)
    if not m:
        raise ValueError("unsupported record format %r" % fmt)
    fields = []
    for part in m.group(1).split(","):
        typ, name = part.split()
        if typ not in _TYPES:
            raise ValueError("unknown field type %r" % typ)
        fields.append((name, _TYPES[typ]))
    return fields


def header_text(fmt, dims, user):
    out = ["<HDR>\n", "%-20s = %s\n" % ("record format", fmt)]
    for i, (n, lbl) in enumerate(dims, 1):
        out.append("%-20s = %-10d %s\n" % ("record dimension %d" % i, n, lbl))
    for i, u in enumerate(user, 1):
        out.append("%-20s =%s\n" % ("user data %d" % i, u))
    out.append("<\\HDR>")
    return "".join(out)


def _parse_header(text):
    fmt = re.search(r"record format\s*=(.*)", text).group(1)
    dims = [(int(n), lbl) for n, lbl in
            re.findall(r"record dimension \d+\s*=\s*(\d+)\s*(.*)", text)]
    user = re.findall(r"user data \d+\s*=(.*)", text)
    return fmt, dims, user


def read(path):
    b = open(path, "rb").read()
    end = b.find(b"<\\HDR>")
    if not b.startswith(b"<HDR>") or end < 0:
        raise ValueError("%s: not an nd file" % path)
    fmt, dims, user = _parse_header(b[:end].decode())
    dt = np.dtype(parse_format(fmt))
    count = int(np.prod([n for n, _ in dims])) if dims else 0
    data = np.frombuffer(b, dtype=dt, count=count, offset=end + 6)
    return NdFile(fmt.strip(), dims, user, data)


def write(path, fmt, dims, user, data):
    dt = np.dtype(parse_format(fmt))
    arr = np.asarray(data)
    if arr.dtype != dt:
        rec = np.empty(arr.shape[0], dtype=dt)
        for k, (name, _) in enumerate(dt.descr):
            rec[name] = arr[:, k]
        arr = rec
    count = int(np.prod([n for n, _ in dims]))
    if arr.shape[0] != count:
        raise ValueError("%d records for dimensions %s" % (arr.shape[0], dims))
    with open(path, "wb") as f:
        f.write(header_text(fmt, dims, user).encode())
        f.write(arr.tobytes())


def to_ascii(path, out=None):
    """The reference's ASCII conversion of a binary nd file (double
    precision): header then one formatted line per record."""
    nd = read(path)
    out = out or path + "_ascii"
    # the converter re-reads the header text after each '=' including the
    # space that follows it and prints it back after "= " / "=": the ASCII
    # header gains one space in the record format and the user data lines
    lines = [header_text(" " + nd.fmt, nd.dims, [" " + u for u in nd.user]), "\n"]
    fmts = []
    for name, dt in parse_format(nd.fmt):
        fmts.append("%21.12e" if dt == "<f8" else "%16.7e" if dt == "<f4" else "%10d")
    names = nd.data.dtype.names
    for rec in nd.data:
        lines.append("".join(f % rec[n] for f, n in zip(fmts, names)) + "\n")
    with open(out, "w") as f:
        f.write("".join(lines))
    return out


# ---- files of the solver path -------------------------------------------

def _common_user(band_model, crystal=2):
    return ["%s  # band_model" % band_model, "%d # crystal structure" % crystal]


def write_ek(path, energies, k=(0.0, 0.0, 0.0)):
    E = np.asarray(energies, dtype=np.float64)
    rec = np.zeros((E.size, 4))
    rec[:, :3] = k
    rec[:, 3] = E
    write(path, "{real kx, real ky, real kz, real E}",
          [(E.size, "# number of eigenvalues")], ["# not used"], rec)


def write_evec(path, psi, nb, band_model):
    psi = np.asarray(psi, dtype=np.complex128).reshape(-1)
    rec = np.stack([psi.real, psi.imag], axis=1)
    write(path, "{real wf.r, real wf.i}",
          [(nb, "# order of orbital basis"), (psi.size // nb, "# number of atoms")],
          _common_user(band_model), rec)


def write_ratom(path, positions, band_model):
    p = np.asarray(positions, dtype=np.float64)
    write(path, "{real x, real y, real z}", [(p.shape[0], "# number of atoms")],
          _common_user(band_model), p)


def write_nbrindx(path, nbr, band_model):
    n = np.asarray(nbr, dtype=np.int32)
    write(path, "{int n1, int n2, int n3, int n4}",
          [(n.shape[0], "# number of atoms")], _common_user(band_model), n)


def write_atype(path, types, band_model):
    t = np.asarray(types, dtype=np.int32).reshape(-1, 1)
    write(path, "{int t}", [(t.shape[0], "# number of atoms")],
          _common_user(band_model), t)


def write_disp(path, disp, band_model):
    d = np.asarray(disp, dtype=np.float64)
    write(path, "{real dx, real dy, real dz}", [(d.shape[0], "# number of atoms")],
          _common_user(band_model), d)


def read_ek(path):
    return np.asarray(read(path).data["E"])


def read_evec(path):
    d = read(path).data
    return d["wf.r"] + 1j * d["wf.i"]


def write_tridiagonal(path, alpha, beta, processors=1):
    """The Lanczos tridiagonal in the reference's binary checkpoint
    (saveTridiagonalMatrix, NEMO_3D/eigen/src/lanczos.c:1692-1714): int32
    iterations, int32 processors, then 2 * iterations doubles alpha_0,
    beta_0, alpha_1, beta_1, ...; the last slot (beta of the final step,
    which the reference leaves unset) holds beta_(n-1)."""
    a = np.asarray(alpha, dtype="<f8")
    b = np.asarray(beta, dtype="<f8")
    n = a.size
    t = np.empty(2 * n, dtype="<f8")
    t[0::2] = a
    t[1::2] = b[:n]
    with open(path, "wb") as f:
        f.write(np.array([n, processors], dtype="<i4").tobytes())
        f.write(t.tobytes())


def read_tridiagonal(path):
    """(alpha, beta, processors) of a checkpoint written by
    write_tridiagonal or the reference (restoreTridiagonalMatrix)."""
    raw = open(path, "rb").read()
    n, procs = np.frombuffer(raw[:8], dtype="<i4")
    t = np.frombuffer(raw[8:8 + 16 * int(n)], dtype="<f8")
    return t[0::2].copy(), t[1::2].copy(), int(procs)

