"""Writes tests/golden/formats/: small files in the reference's on-disk formats and
what the reference's OWN readers make of them (oracle/_ref/ref_driver, built by
oracle/ref.mk from the reference's sources in this container).

  * par_*.par: two of the reference's sample parameter files (samples/reconMLEM.par,
    samples/cumlem/cumlem.par, copied as data) and seeded synthetic files that mix
    the parser's corner cases (leading blanks and tabs, "key:= v", "key :=v",
    comments, lines without a separator, CR / CRLF / LF line ends, a line longer than
    the parser keeps, a missing final line end, wrong first lines, duplicates).
    Keys stay below the parser's 256-character key buffer and long lines put the
    separator late enough that the value fits its buffer: beyond those limits the
    reference's parser overruns its stack buffers, and no answer is defined.
  * itf_*.h33: seeded synthetic Interfile headers for interfile_read.
  * par_expected.json / itf_expected.json: parametersFile_read / interfile_read
    return code and value for every (file, method, key) probed.
  * log_row.json: Logger::writeRowOfNumbers (the log's number rows) on a few values.
  * ref_image.h33/.i33, ref_sino3d.h33/.i33, ref_sino2d.h33/.i33: written by the
    reference's Image / Sinogram3D / Sinogram2D writers; formats.npz holds the
    arrays they were given and the geometry.

Run from the repo root:  python tests/golden/make_format_goldens.py"""
import json
import os
import shutil
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
OUT = os.path.join(HERE, "formats")
DRIVER = os.path.join(ROOT, "oracle", "_ref", "ref_driver")
REF_SAMPLES = "/root/reference/samples"

PAR_KEYS = ["input type", "input file", "initial estimate", "output filename prefix",
            "number of iterations", "number of subsets", "save estimates at iteration intervals",
            "forwardprojector", "backprojector", "number of iteration", "a", "", ";a", "END",
            "k" * 254]
ITF_KEYS = ["!matrix size [1]", "number of rings", "!name of data file", "x", "", " x",
            "!END OF INTERFILE"]
SINO3D = dict(nx=8, ny=8, nz=7, dx=4.0, dz=4.0, nproj=12, nr=20, rings=4, rfov=32.0, zfov=28.0,
              rscanner=120.0, span=3, mrd=2)


def driver(*args):
    out = subprocess.run([DRIVER] + list(args), check=True, capture_output=True)
    return out.stdout


def lookup(cmd, path, key, method=None):
    if cmd == "par_read":
        raw = driver("par_read", f"file={path}", f"method={method}", f"key={key}")
    else:
        raw = driver("itf_read", f"header={path}", f"key={key}")
    rc, _, value = raw.rstrip(b"\n").partition(b"\t")
    return [int(rc), value.decode("latin-1")]


def par_text(rng):
    """One synthetic parameter file (bytes)."""
    keys = PAR_KEYS[:9] + ["a", "number of iterations"]
    vals = ["12", "sino.h33", " 7", "", "a b", "3.5", "Siddon"]
    ends = [b"\n"] * 7 + [b"\r\n"] * 2 + [b"\r"]
    first = [b"MLEM Parameters :=", b"OSEM Parameters :=", b"MLEM Parameters:=",
             b"  OSEM Parameters :="][int(rng.choice(4, p=[0.45, 0.45, 0.05, 0.05]))]
    lines = [first, b"number of subsets := 4"]  # a plain first entry: the parser's
    # key start is carried over from the previous line for an all-blank key
    for _ in range(int(rng.integers(4, 14))):
        k = keys[int(rng.integers(len(keys)))].encode()
        v = vals[int(rng.integers(len(vals)))].encode()
        kind = int(rng.integers(14))
        lines.append([k + b" := " + v, b"  " + k + b" := " + v, b"\t" + k + b" := " + v,
                      k + b":= " + v, k + b" :=" + v, k + b" :=  " + v + b"  ", b";" + k + b" := " + v,
                      b" ;a := " + v, k, b"", b"   := x", k + b" := " + v + b" := w", b"END :=",
                      b"k" * 254 + b" := " + b"v" * 400][kind])
    body = b""
    for ln in lines:
        body += ln + ends[int(rng.integers(len(ends)))]
    if rng.random() < 0.15:
        body = body.rstrip(b"\r\n")
    return body


def itf_text(rng):
    keys = [b"!matrix size [1]", b"number of rings", b"!name of data file", b"x", b" x"]
    vals = [b"12", b"data.i33", b" 7", b"", b"{1,2}"]
    first = b"!INTERFILE :=" if rng.random() < 0.9 else b"INTERFILE :="
    lines = [first]
    for _ in range(int(rng.integers(4, 12))):
        k = keys[int(rng.integers(len(keys)))]
        v = vals[int(rng.integers(len(vals)))]
        kind = int(rng.integers(9))
        lines.append([k + b" := " + v, b"  " + k + b" := " + v, k + b":= " + v, b";" + k + b" := " + v,
                      k, b"", b"!END OF INTERFILE :=", k + b" := " + v + b"\r",
                      b"y" * 600 + b" := " + v][kind])
    body = b"\n".join(lines) + b"\n"
    if rng.random() < 0.15:
        body = body.rstrip(b"\n")
    return body


def main():
    if not os.path.exists(DRIVER):
        sys.exit(f"{DRIVER} is not built: make -f oracle/ref.mk")
    os.makedirs(OUT, exist_ok=True)
    shutil.copy(os.path.join(REF_SAMPLES, "reconMLEM.par"), os.path.join(OUT, "par_reconMLEM.par"))
    shutil.copy(os.path.join(REF_SAMPLES, "cumlem", "cumlem.par"), os.path.join(OUT, "par_cumlem.par"))
    rng = np.random.default_rng(20261202)
    for k in range(24):
        with open(os.path.join(OUT, f"par_synth{k:02d}.par"), "wb") as f:
            f.write(par_text(rng))
    with open(os.path.join(OUT, "par_empty.par"), "wb") as f:
        pass
    for k in range(12):
        with open(os.path.join(OUT, f"itf_synth{k:02d}.h33"), "wb") as f:
            f.write(itf_text(rng))

    par = {}
    for name in sorted(x for x in os.listdir(OUT) if x.startswith("par_") and x.endswith(".par")):
        par[name] = {m: {key: lookup("par_read", os.path.join(OUT, name), key, m) for key in PAR_KEYS}
                     for m in ("MLEM", "OSEM")}
    itf = {}
    for name in sorted(x for x in os.listdir(OUT) if x.startswith("itf_") and x.endswith(".h33")):
        itf[name] = {key: lookup("itf_read", os.path.join(OUT, name), key) for key in ITF_KEYS}

    # files written by the reference's writers, with relative names (run in OUT)
    img = np.random.default_rng(1).random((3, 6, 8), dtype=np.float32) * 5
    s3 = SINO3D
    sino3 = np.random.default_rng(2).random((11, s3["nproj"], s3["nr"]), dtype=np.float32)
    sino2 = np.random.default_rng(3).random((12, 20), dtype=np.float32)
    img.tofile(os.path.join(OUT, "_img.f32"))
    sino3.tofile(os.path.join(OUT, "_s3.f32"))
    sino2.tofile(os.path.join(OUT, "_s2.f32"))
    cwd = os.getcwd()
    os.chdir(OUT)
    try:
        driver("img_write", "nx=8", "ny=6", "nz=3", "dx=2.5", "dz=3.25", "in=_img.f32",
               "prefix=ref_image")
        driver("sino3d_write", *[f"{k}={v}" for k, v in s3.items()], "in=_s3.f32",
               "header=ref_sino3d")
        driver("sino2d_write", "nproj=12", "nr=20", "rfov=32", "rscanner=120", "in=_s2.f32",
               "prefix=ref_sino2d")
    finally:
        os.chdir(cwd)
    for f in ("_img.f32", "_s3.f32", "_s2.f32"):
        os.remove(os.path.join(OUT, f))
    # Logger::writeRowOfNumbers on values of every printing regime
    row = np.array([1234567.0, 0.0001, 100.0, -2.5, 3.14159265, 1e-7, 6.02e23, 0.0, -0.0,
                    123456.75, 1.5e-5, 99999.95, 16777217.0, -3244243.8, 446506.9],
                   np.float32)
    row.tofile(os.path.join(OUT, "_row.f32"))
    driver("log_row", f"n={row.size}", f"in={os.path.join(OUT, '_row.f32')}",
           f"out={os.path.join(OUT, '_row.log')}")
    with open(os.path.join(OUT, "_row.log"), "rb") as f:
        row_text = f.read().decode("latin-1")
    os.remove(os.path.join(OUT, "_row.f32"))
    os.remove(os.path.join(OUT, "_row.log"))
    with open(os.path.join(OUT, "log_row.json"), "w") as f:
        json.dump({"values": [float(v) for v in row], "text": row_text}, f, indent=0)
    np.savez_compressed(os.path.join(OUT, "formats.npz"), image=img, image_voxel=np.float32([2.5, 2.5, 3.25]),
                        sino3d=sino3, sino3d_geom=np.array(json.dumps(s3)), sino2d=sino2)
    with open(os.path.join(OUT, "par_expected.json"), "w") as f:
        json.dump(par, f, indent=0, sort_keys=True)
    with open(os.path.join(OUT, "itf_expected.json"), "w") as f:
        json.dump(itf, f, indent=0, sort_keys=True)
    print(f"wrote {len(par)} parameter files, {len(itf)} headers, 3 reference-written files")


if __name__ == "__main__":
    main()
