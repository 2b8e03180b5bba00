"""The reference's on-disk formats: .par parameter files and Interfile images and
sinograms.

Parameter files (data/src/ParametersFile.cpp:56-227, parametersFile_read), read
byte by byte as the reference reads them:
  * The first line (up to '\\r' or '\\n') must start with "<method> Parameters :=",
    else the lookup fails with code 5; an empty file fails with code 4.
  * Lines end at '\\r' or '\\n' (a CRLF pair leaves an empty line between) and hold at
    most 516 characters; the character after a full line is dropped.
  * A line whose first character is ';' is a comment.
  * The separator is the first ":=" in characters 0..515 (not found: the whole line
    is the key).  The key runs from the first character that is neither blank nor
    tab, within the first pos-2 characters, to two characters before the '=' -- so
    "key := value" gives "key", "key:= v" gives "ke" -- and holds at most 256
    characters; when every candidate is blank the previous line's start is reused.
    The value runs from two characters after the '=' to the end of the line.
  * A file whose last line has no line end fails with code 4, whatever the key.
  * A key found twice fails with code 1 (the last value is returned); not found, 2.
  * The "END :=" test compares the parsed key, which never holds ":=", so the whole
    file is read.

Interfile headers (data/src/Interfile.cpp:70-150, interfile_read), used for 3D
sinograms: the same split with no trimming of leading blanks, lines end at '\\n'
only (a '\\r' ends the value), the first line must start with "!INTERFILE" (else
code 4), and reading stops at the first key that starts with "!END OF INTERFILE".

Images and 2D sinograms (data/src/Images.cpp:156-381, Sinogram2D.cpp:319-381) go
through the vendored medcon Interfile module, which matches keys case-insensitively
with blanks removed and strips comments after ';' (data/medcon/m-intf.c:150-232).
The reader here applies those rules to the keys that fix a float image's size,
voxel size, byte order and data file; the writer emits the keys that module reads.
A 2D sinogram is a one-image file of width nr and height nproj.  3D sinograms
(Sinogram3D.cpp:106-471, writer :609-750) store segments, then sinograms, then views,
then radial bins, with the segment table in the header."""
import os
import re

import numpy as np

PAR_LINE = 516       # ParametersFile.cpp: characters kept per line
PAR_KEY = 256        # ... per key
PAR_VALUE_END = 512  # value characters end before this line index
ITF_LINE = 1024      # Interfile.cpp: characters kept per line
ITF_SEP = 512        # ... the separator is searched for below this index


class FormatError(RuntimeError):
    def __init__(self, msg, code=None):
        super().__init__(msg)
        self.code = code


def atoi(text):
    """C atoi: leading white space, an optional sign, then digits; 0 if none."""
    m = re.match(r"[ \t\n\r\f\v]*([+-]?\d+)", text)
    return int(m.group(1)) if m else 0


def atof(text):
    """C atof for decimal input: the longest leading number, 0.0 if none."""
    m = re.match(r"[ \t\n\r\f\v]*([+-]?(\d+\.?\d*|\.\d+)([eE][+-]?\d+)?)", text)
    return float(m.group(1)) if m else 0.0


def _cstr(b):
    k = b.find(0)
    return bytes(b if k < 0 else b[:k])


def _first_line(data):
    """Bytes of the first line up to and including its '\\r' or '\\n', or None when the
    file ends first."""
    for k, c in enumerate(data):
        if c in (10, 13):
            return data[:k + 1], k + 1
    return None, len(data)


def _scan_par(data, method):
    """(code, [(key, value)]) by parametersFile_read's reading of the whole file."""
    if not data:
        return 4, []
    first, p = _first_line(data)
    if first is None:
        return 4, []
    head = f"{method} Parameters :=".encode("latin-1")
    if first[:len(head)] != head:
        return 5, []
    entries, pos0, n = [], 0, len(data)
    while p < n:
        c = data[p]
        p += 1
        line = bytearray(520)
        i = 0
        while c != 13 and c != 10 and i < PAR_LINE:
            line[i] = c
            if p >= n:
                return 4, entries
            c = data[p]
            p += 1
            i += 1
        if line[0] == 59:  # ';'
            continue
        pos = PAR_LINE
        for q in range(1, PAR_LINE):
            if line[q] == 61 and line[q - 1] == 58:
                pos = q
                break
        for q in range(0, min(pos - 2, PAR_KEY)):
            if line[q] not in (9, 32):
                pos0 = q
                break
        key = _cstr(line[pos0:max(pos0, min(pos - 2, PAR_KEY))])
        val = bytearray()
        for q in range(pos + 2, min(PAR_KEY + pos + 2, PAR_VALUE_END)):
            if line[q] in (0, 13, 10):
                break
            val.append(line[q])
        if key[:6] == b"END :=":
            break
        entries.append((key, bytes(val)))
    return 0, entries


def _scan_interfile(data):
    """(code, [(key, value)]) by interfile_read's reading of the whole header."""
    first, p = _first_line(data)
    if first is None or first[:10] != b"!INTERFILE":
        return 4, []
    entries, n = [], len(data)
    while p < n:
        c = data[p]
        p += 1
        line = bytearray(ITF_LINE + 4)
        i = 0
        while c != 10 and i < ITF_LINE:
            line[i] = c
            if p >= n:
                return 4, entries
            c = data[p]
            p += 1
            i += 1
        if line[0] == 59:
            continue
        pos = ITF_SEP
        for q in range(1, ITF_SEP):
            if line[q] == 61 and line[q - 1] == 58:
                pos = q
                break
        key = _cstr(line[0:max(0, pos - 2)])
        val = bytearray()
        for q in range(pos + 2, ITF_LINE):
            if line[q] in (0, 13, 10):
                break
            val.append(line[q])
        if key[:17] == b"!END OF INTERFILE":
            break
        entries.append((key, bytes(val)))
    return 0, entries


class _Keyed:
    """Lookups over one scan with the reference's return codes:
    0 found once, 1 found more than once (last value), 2 not found, and the scan's
    own code (3 unreadable, 4 format, 5 wrong method) for every key."""

    def __init__(self, path, code, entries):
        self.path, self.code, self.entries = path, code, entries

    def read(self, key):
        """(code, value): with code 4 after a partial scan, the value is the last one
        found before the scan stopped, as the reference leaves it."""
        kb = key.encode("latin-1")
        hits = [v for k, v in self.entries if k == kb]
        last = hits[-1].decode("latin-1") if hits else ""
        if self.code:
            return self.code, last
        if not hits:
            return 2, ""
        return (0 if len(hits) == 1 else 1), last

    def get(self, key, default=None, required=False):
        """The value, `default` when the key is absent and not `required`; any other
        outcome raises FormatError carrying the reference's code."""
        code, value = self.read(key)
        if code == 0:
            return value
        if code == 2 and not required:
            return default
        why = {1: f"keyword '{key}' appears more than once",
               2: f"keyword '{key}' not found",
               3: "could not be opened for reading",
               4: "wrong file format",
               5: "first line is not the expected 'Method Parameters :='"}[code]
        raise FormatError(f"{self.path}: {why}", code)


def _read_bytes(path):
    try:
        with open(path, "rb") as f:
            return f.read(), 0
    except OSError:
        return b"", 3


class ParFile(_Keyed):
    """A .par file of `method` ("OSEM", "MLEM", ...)."""

    def __init__(self, path, method):
        data, code = _read_bytes(path)
        if code == 0:
            code, entries = _scan_par(data, method)
        else:
            entries = []
        super().__init__(path, code, entries)
        self.method = method


class Header(_Keyed):
    """An Interfile header read as interfile_read reads it."""

    def __init__(self, path):
        data, code = _read_bytes(path)
        if code == 0:
            code, entries = _scan_interfile(data)
        else:
            entries = []
        super().__init__(path, code, entries)


# ---- medcon-style images ---------------------------------------------------------


def _medcon_keys(path):
    """[(key, value)] with medcon's normalisation: text after ';' dropped, keys
    lower-cased without blanks and without a leading '!', values stripped."""
    with open(path, "rb") as f:
        text = f.read().decode("latin-1")
    lines = text.replace("\r", "\n").split("\n")
    if not lines or "interfile" not in lines[0].lower().replace(" ", ""):
        raise FormatError(f"{path}: not an Interfile header")
    out = []
    for raw in lines[1:]:
        line = raw.split(";", 1)[0]
        if ":=" not in line:
            continue
        k, v = line.split(":=", 1)
        k = re.sub(r"\s+", "", k).lower().lstrip("!")
        out.append((k, v.strip()))
    return out


def _find(keys, name, default=None):
    """Last value whose key contains `name` (medcon matches substrings)."""
    want = re.sub(r"\s+", "", name).lower()
    hit = default
    for k, v in keys:
        if want in k:
            hit = v
    return hit


def _data_file(header_path, name):
    """The data file's path: as written when it holds a directory, else beside the
    header (m-intf.c:640-660, Sinogram3D.cpp:115-130)."""
    if "/" in name:
        return name
    return os.path.join(os.path.dirname(header_path), name)


def read_image(header_path):
    """Interfile image (or 2D sinogram) -> (float32 array [nz, ny, nx], (dx, dy, dz))."""
    keys = _medcon_keys(header_path)
    fmt = (_find(keys, "number format", "short float") or "").lower()
    if fmt not in ("short float", "float") or atoi(_find(keys, "number of bytes per pixel",
                                                         "4")) != 4:
        raise FormatError(f"{header_path}: only 4-byte float data is supported")
    nx = atoi(_find(keys, "matrix size [1]", "0"))
    ny = atoi(_find(keys, "matrix size [2]", "0"))
    nz = _find(keys, "matrix size [3]")
    nz = atoi(nz) if nz is not None else atoi(_find(keys, "total number of images", "1"))
    if nx <= 0 or ny <= 0 or nz <= 0:
        raise FormatError(f"{header_path}: matrix sizes are missing")
    dx = atof(_find(keys, "scaling factor (mm/pixel) [1]", "1"))
    dy = atof(_find(keys, "scaling factor (mm/pixel) [2]", str(dx)))
    dz = atof(_find(keys, "scaling factor (mm/pixel) [3]", "1"))
    order = (_find(keys, "imagedata byte order", "littleendian") or "").lower()
    dt = np.dtype(">f4") if order.startswith("big") else np.dtype("<f4")
    off = atoi(_find(keys, "data offset in bytes", "0") or "0")
    name = _find(keys, "name of data file")
    if not name:
        raise FormatError(f"{header_path}: no data file named")
    data = np.fromfile(_data_file(header_path, name), dtype=dt, count=nx * ny * nz,
                       offset=off)
    if data.size != nx * ny * nz:
        raise FormatError(f"{header_path}: data file holds fewer than {nx * ny * nz} values")
    return data.astype(np.float32).reshape(nz, ny, nx), (dx, dy, dz)


def write_image(prefix, img, dx, dy=None, dz=1.0):
    """Writes prefix.h33 / prefix.i33 (float, little endian) readable by the
    reference's Image::readFromInterfile; returns the header's path."""
    img = np.ascontiguousarray(img, np.float32)
    if img.ndim == 2:
        img = img[None]
    nz, ny, nx = img.shape
    dy = dx if dy is None else dy
    img.astype("<f4").tofile(prefix + ".i33")
    lines = [
        "!INTERFILE :=",
        "!imaging modality := nucmed",
        "!version of keys := 3.3",
        "!GENERAL DATA :=",
        "!data offset in bytes := 0",
        f"!name of data file := {os.path.basename(prefix)}.i33",
        "!GENERAL IMAGE DATA :=",
        "!type of data := pet",
        f"!total number of images := {nz}",
        "imagedata byte order := LITTLEENDIAN",
        "number of energy windows := 1",
        "!PET STUDY (General) :=",
        "!process status := Reconstructed",
        f"number of images/energy window := {nz}",
        f"!matrix size [1] := {nx}",
        "matrix axis label [1] := x",
        f"scaling factor (mm/pixel) [1] := {dx:.6f}",
        f"!matrix size [2] := {ny}",
        "matrix axis label [2] := y",
        f"scaling factor (mm/pixel) [2] := {dy:.6f}",
        f"!matrix size [3] := {nz}",
        "matrix axis label [3] := z",
        f"scaling factor (mm/pixel) [3] := {dz:.6f}",
        "!number format := short float",
        "!number of bytes per pixel := 4",
        "!SPECT STUDY (reconstructed data) :=",
        f"!number of slices := {nz}",
        "!END OF INTERFILE :=",
    ]
    with open(prefix + ".h33", "w") as f:
        f.write("\n".join(lines) + "\n")
    return prefix + ".h33"


def read_sinogram2d(header_path):
    """2D sinogram -> float32 [nproj, nr] (Sinogram2D::readFromInterfile: one image)."""
    s, _ = read_image(header_path)
    if s.shape[0] != 1:
        raise FormatError(f"{header_path}: a 2D sinogram holds one image")
    return s[0]


def write_sinogram2d(prefix, sino):
    """Writes a 2D sinogram [nproj, nr] as a one-image Interfile file."""
    sino = np.ascontiguousarray(sino, np.float32)
    return write_image(prefix, sino[None], 1.0, 1.0, 1.0)


# ---- 3D sinograms ----------------------------------------------------------------


def _brace_list(text, n, seps, what):
    """Sinogram3D.cpp's parse of "{a,b,...}": between the last '{' and the last '}',
    tokens split on `seps` (the first on " ,"), exactly n of them, each by atoi."""
    a, b = text.rfind("{"), text.rfind("}")
    if a < 0 or b < 0:
        raise FormatError(f"no {{}} around the {what} list")
    body = text[a + 1:b] if b > a else ""
    # strtok: the first token ends at ' ' or ',', later ones at any of `seps`; runs
    # of separators count as one
    m = re.match(r"[ ,]*([^ ,]+)", body)
    if not m:
        raise FormatError(f"the {what} list is empty")
    pat = "[" + re.escape(seps) + "]"
    toks = [m.group(1)] + [t for t in re.split(pat, body[m.end():]) if t]
    if len(toks) < n:
        raise FormatError(f"the {what} list has fewer elements than segments")
    if len(toks) > n:
        raise FormatError(f"the {what} list has more elements than segments")
    return [atoi(t) for t in toks]


def segment_rings(rings, minrd, maxrd):
    """Per segment, the ring pairs of each of its sinograms in file order, by the
    reference's enumeration (Sinogram3D.cpp:384-441): z1 runs over 0..2*rings-1 and a
    sinogram holds every in-range (z1 - j, j) with ring difference in [min, max]."""
    out = []
    for mn, mx in zip(minrd, maxrd):
        sinos = []
        for z1 in range(2 * rings):
            pairs = []
            for z2 in range(rings):
                a = z1 - z2
                if mn <= a - z2 <= mx and 0 <= a < rings:
                    pairs.append((a, z2))
            if pairs:
                sinos.append(pairs)
        out.append(sinos)
    return out


def read_sinogram3d(header_path):
    """3D sinogram -> dict(bins [nsino, nproj, nr], nsino, minrd, maxrd, rings, nproj,
    nr) with the checks of Sinogram3D::readFromInterfile.  As there, a data file that
    ends on a sinogram boundary leaves the remaining sinograms zero, and one that ends
    inside a sinogram is an error."""
    h = Header(header_path)
    name = _data_file(header_path, h.get("!name of data file", required=True))
    if atoi(h.get("number of dimensions", required=True)) != 4:
        raise FormatError(f"{header_path}: a 3D sinogram has four dimensions")

    def label(i, want):
        if h.get(f"matrix axis label [{i}]", required=True) != want:
            raise FormatError(f"{header_path}: dimension {i} must be '{want}'")

    label(4, "segment")
    nseg = atoi(h.get("!matrix size [4]", required=True))
    if nseg <= 0:
        raise FormatError(f"{header_path}: the sinogram has no segments")
    label(3, "axial coordinate")
    nsino = _brace_list(h.get("!matrix size [3]", required=True), nseg, " ,.-", "sinograms")
    minrd = _brace_list(h.get("minimum ring difference per segment", required=True), nseg,
                        " ,", "minimum ring difference")
    maxrd = _brace_list(h.get("maximum ring difference per segment", required=True), nseg,
                        " ,", "maximum ring difference")
    label(2, "view")
    nproj = atoi(h.get("!matrix size [2]", required=True))
    label(1, "tangential coordinate")
    nr = atoi(h.get("!matrix size [1]", required=True))
    code, v = h.read("number of rings")
    rings = atoi(v) if code == 0 else (nsino[0] + 1) // 2
    layout = segment_rings(rings, minrd, maxrd)
    for s, sinos in enumerate(layout):
        if len(sinos) != nsino[s]:
            raise FormatError(f"{header_path}: segment {s} has {len(sinos)} sinograms for "
                              f"its ring range, the header says {nsino[s]}")
    total, per = sum(nsino), nproj * nr
    try:
        raw = np.fromfile(name, dtype="<f4", count=total * per)
    except OSError as e:
        raise FormatError(f"{header_path}: cannot read the data file {name}: {e}") from e
    whole = raw.size // per
    if raw.size % per and whole < total:
        raise FormatError(f"{header_path}: the data file ends inside sinogram {whole}")
    bins = np.zeros((total, nproj, nr), np.float32)
    bins.reshape(-1)[:whole * per] = raw[:whole * per]
    return dict(bins=bins, nsino=nsino, minrd=minrd, maxrd=maxrd, rings=rings, nproj=nproj,
                nr=nr)


def write_sinogram3d(prefix, bins, nsino, minrd, maxrd, rings, dr_mm, dang_deg):
    """Writes prefix.h33 / prefix.i33 with the keys of Sinogram3D::writeInterfile;
    returns the header's path."""
    bins = np.ascontiguousarray(bins, np.float32)
    total, nproj, nr = bins.shape
    if total != sum(nsino):
        raise FormatError("bins do not match the segment table")
    bins.astype("<f4").tofile(prefix + ".i33")
    lst = lambda v: "{" + ",".join(str(int(x)) for x in v) + "}"  # noqa: E731
    lines = [
        "!INTERFILE :=",
        "!imaging modality := nucmed",
        "!version of keys := 3.3",
        "!GENERAL DATA := ",
        "!data starting block := 0",
        f"!name of data file := {os.path.basename(prefix)}.i33",
        "!GENERAL IMAGE DATA :=",
        "!type of data := PET",
        "!imagedata byte order := LITTLEENDIAN",
        "!PET STUDY (General) :=",
        "!number format := short float",
        "!number of bytes per pixel := 4",
        f"scaling factor (mm/pixel) [1] := {dr_mm:g}",
        f"scaling factor (deg/pixel) [2] := {dang_deg:g}",
        "number of dimensions := 4",
        "matrix axis label [4] := segment",
        f"!matrix size [4] := {len(nsino)}",
        "matrix axis label [2] := view",
        f"!matrix size [2] := {nproj}",
        "matrix axis label [1] := tangential coordinate",
        f"!matrix size [1] := {nr}",
        "matrix axis label [3] := axial coordinate",
        f"!matrix size [3] := {lst(nsino)}",
        f"minimum ring difference per segment := {lst(minrd)}",
        f"maximum ring difference per segment := {lst(maxrd)}",
        f"number of rings := {rings}",
        "!data offset in bytes[1] := 0",
        "!END OF INTERFILE :=",
    ]
    with open(prefix + ".h33", "w") as f:
        f.write("\n".join(lines) + "\n")
    return prefix + ".h33"
