"""The command-line drop-in's handling of parameter files, without a GPU: every
refusal happens before a device is touched.  The keys and their rules follow
cmd/src/OSEM.cpp:220-300 and data/src/readParameters.cpp."""
import numpy as np
import pytest

from apirl_amd import cli, interfile

BASE = """{method} Parameters :=
input type := Sinogram3D
input file := sino.h33
initial estimate := {init}
output filename prefix := {out}
number of iterations := 2
number of subsets := 4
forwardprojector := Siddon
backprojector := Siddon
cylindrical pet radius (in mm) := 300
radius fov (in mm) := 64
axial fov (in mm) := 36
"""


@pytest.fixture
def files(tmp_path):
    init = interfile.write_image(str(tmp_path / "init"), np.ones((9, 32, 32), np.float32), 4.0,
                                 4.0, 4.0)

    def make(extra="", drop=(), method="OSEM", name="r.par"):
        text = BASE.format(method=method, init=init, out=tmp_path / "out")
        lines = [ln for ln in text.splitlines() if not any(ln.startswith(d) for d in drop)]
        path = tmp_path / name
        path.write_text("\n".join(lines) + "\n" + extra)
        return str(path)

    return make


def refused(path, match, method="OSEM"):
    with pytest.raises(interfile.FormatError, match=match):
        cli.run(method, path, log=lambda *_: None)


def test_extension_required(files):
    refused(files(name="r.txt"), "extension")


def test_first_line_must_name_the_method(files):
    refused(files(method="MLEM"), "first line")


@pytest.mark.parametrize("key", ["input type", "input file", "initial estimate",
                                 "output filename prefix", "number of iterations",
                                 "number of subsets", "forwardprojector", "backprojector"])
def test_required_keys(files, key):
    refused(files(drop=(key,)), f"'{key}' not found")


def test_duplicate_key(files):
    refused(files(extra="number of iterations := 3\n"), "more than once")


@pytest.mark.parametrize("extra,match", [
    ("siddon number of samples on the detector := 0\n", "at least one sample"),
    ("siddon number of axial samples on the detector := x\n", "at least one sample"),
    ("attenuation image filename := mu.h33\n", "not available on this path"),
])
def test_unsupported_options_refused(files, extra, match):
    refused(files(extra=extra), match)


def test_unknown_projector(files):
    path = files(drop=("forwardprojector",), extra="forwardprojector := RotationBasedProjector\n")
    refused(path, "RotationBasedProjector")


def test_main_reports_errors(files, capsys):
    assert cli.main(["OSEM", files(drop=("backprojector",))]) == 1
    assert "backprojector" in capsys.readouterr().err
    assert cli.main(["FBP", "x.par"]) == 1


def test_log_rows_match_reference_logger():
    """Number rows of the log as Logger::writeRowOfNumbers prints them (golden from the
    reference's Logger, tests/golden/formats/log_row.json)."""
    import json
    import os
    path = os.path.join(os.path.dirname(__file__), "golden", "formats", "log_row.json")
    with open(path) as f:
        g = json.load(f)
    assert cli._row(g["values"]) + "\r\n" == g["text"]


def test_log_layout(tmp_path):
    prefix = str(tmp_path / "osem")
    times = np.arange(8, dtype=np.float32).reshape(4, 2)
    cli.write_log(prefix, [("Iterations", "1"), ("Rings", "4")], [0], 12.5, times,
                  [1.5, 2.5, 3.5])
    text = open(prefix + ".log", "rb").read().decode("utf-8")
    lines = text.split("\r\n")
    assert lines[0] == "######## ML-EM Reconstruction #########"
    assert lines[1] == f"Name: {prefix}" and lines[2] == "Type: ML-EM"
    assert lines[3:5] == ["Iterations: 1", "Rings: 4"]
    assert lines[5] == "_______RECONSTRUCCION_______"
    assert lines[6] == f"Imagen de iteración 0 guardada en: {prefix}_iter_0"
    assert lines[7] == f"Imagen final guardada en: {prefix}_final"
    assert lines[9] == "Tiempo Total de Reconstrucción:: 12.500000"
    assert lines[10:12] == ["Tiempos de Reconstrucción por Iteración [mseg]", "0, 1"]
    assert lines[16:18] == ["Tiempos de UpdatePixel por Iteración [mseg]", "6, 7"]
    assert lines[18:20] == ["Likelihood por Iteración:", "1.5, 2.5, 3.5"]
    assert text.endswith("\r\n") and "\n" not in text.replace("\r\n", "")


def test_gpu_block_size_keys(tmp_path):
    """ProcessBlockSizeString (readCudaParameters.cpp:23-49): the text between the
    last '{' and '}' split at blanks and commas; missing braces or elements refused."""
    path = tmp_path / "m.par"
    path.write_text("MLEM Parameters :=\nprojector block size := {128, 2,1}\n"
                    "backprojector block size := 64,1,1\npixel update block size := {8,1}\n")
    par = interfile.ParFile(str(path), "MLEM")
    assert cli._block_size(par, "projector block size") == (128, 2, 1)
    with pytest.raises(interfile.FormatError, match="no {}"):
        cli._block_size(par, "backprojector block size")
    with pytest.raises(interfile.FormatError, match="x, y and z"):
        cli._block_size(par, "pixel update block size")
    with pytest.raises(interfile.FormatError, match="not found"):
        cli._block_size(par, "gpu id")


def test_subset_keys_read_together(tmp_path):
    """"number of subsets" and "subset index" count only together
    (project.cpp:226-237); an index outside the subsets is refused."""
    def par(text):
        path = tmp_path / "p.par"
        path.write_text("Projection Parameters :=\n" + text)
        return interfile.ParFile(str(path), "Projection")
    assert cli._subset_keys(par("number of subsets := 7\nsubset index := 2\n")) == (7, 2)
    assert cli._subset_keys(par("number of subsets := 7\n")) == (0, 0)
    with pytest.raises(interfile.FormatError, match="subset index"):
        cli._subset_keys(par("number of subsets := 7\nsubset index := 7\n"))


RUN_KEYS = {"input type": "Sinogram3D", "input file": "s.h33", "initial estimate": "no_init.h33",
            "output filename prefix": "o", "number of iterations": "1",
            "backprojector": "Siddon"}


@pytest.mark.parametrize("cmd,method,keys,match", [
    ("project", "Projection", {"input file": "i.h33", "output type": "Sinogram3D",
                               "projector": "ConeOfResponse", "output filename": "o",
                               "output projection": "s.h33"}, "ConeOfResponse"),
    ("backproject", "Backproject", {"input type": "Sinogram3D", "input file": "s.h33",
                                    "backprojector": "Siddon", "output image": "i.h33"},
     "output filename"),
    ("cuMLEM", "MLEM", {"input type": "Sinogram3D", "projector block size": "{1,1,1}",
                        "backprojector block size": "{1,1,1}"}, "pixel update block size"),
    # the CPU command knows no GPU projector name (readParameters.cpp:117-158) ...
    ("MLEM", "MLEM", dict(RUN_KEYS, forwardprojector="CuSiddonProjector"), "CuSiddonProjector"),
    # ... cuMLEM does (cuMLEM.cpp:279-316): it goes on to read the initial estimate
    ("cuMLEM", "MLEM", dict(RUN_KEYS, forwardprojector="CuSiddonProjector",
                            backprojector="CuSiddonProjector",
                            **{"projector block size": "{1,1,1}",
                               "backprojector block size": "{1,1,1}",
                               "pixel update block size": "{1,1,1}"}), "no_init.h33"),
])
def test_projection_commands_refuse_before_the_device(tmp_path, capsys, cmd, method, keys, match):
    path = tmp_path / "c.par"
    path.write_text(f"{method} Parameters :=\n" +
                    "".join(f"{k} := {v}\n" for k, v in keys.items()))
    assert cli.main([cmd, str(path)]) == 1
    assert match in capsys.readouterr().err
