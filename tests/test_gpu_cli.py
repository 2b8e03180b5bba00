"""The command-line drop-in end to end: Interfile inputs and a .par file in, the
reference's output files out, on the GPU path.

  * 2D MLEM (Sinogram2D input, 60 iterations): the final image and the sensitivity
    equal the reference's Mlem2d output (tests/golden/c1_mlem2d.npz) bit for bit.
  * 3D OSEM (Sinogram3D input, 7 subsets x 2 iterations, saved every iteration):
    the final image is within the image bar (1e-4 relative L2) of the reference's
    OsemSinogram3d output (mid3d_osem.npz), its NaN voxels coincide, and the saved
    iterations, final image and subset sensitivities equal, bit for bit, what
    Plan.reconstruct / Plan.sensitivity give for the same inputs.
  * 3D OSEM with multiplicative and additive sinograms read from Interfile files:
    bit-identical to Plan.reconstruct with the same corrections."""
import json
import os

import numpy as np
import pytest

from apirl_amd import cli, hip, interfile
from test_gpu_parity import (IMG_RL2, assert_bitexact, assert_voxel_close, c1_plan, ctx,  # noqa: F401
                             dev, host, log, plan3d, rel_l2, torch)

pytestmark = pytest.mark.gpu


def write_par(path, method, **keys):
    lines = [f"{method} Parameters :="] + [f"{k} := {v}" for k, v in keys.items()]
    path.write_text("\n".join(lines) + "\n")
    return str(path)


def read(prefix):
    return interfile.read_image(prefix + ".h33")[0]


def log_rows(prefix):
    """{title: numbers} of the number rows of <prefix>.log (reference log format)."""
    lines = open(prefix + ".log", "rb").read().decode("utf-8").split("\r\n")
    rows = {}
    for k, line in enumerate(lines[:-1]):
        if line.startswith("Tiempos de") or line.startswith("Likelihood por"):
            rows[line] = [float(v) for v in lines[k + 1].split(", ")]
    return rows


def test_cli_mlem2d_equals_reference(torch, ctx, golden, tmp_path):
    log("case: CLI MLEM, Sinogram2D input, c1 geometry, 60 iterations")
    f = golden("c1_mlem2d.npz")
    g = json.loads(str(f["geom"]))
    sino = interfile.write_sinogram2d(str(tmp_path / "sino"), f["meas"].astype(np.float32))
    init = interfile.write_image(str(tmp_path / "init"), np.ones((1, 128, 128), np.float32),
                                 g["dx"], g["dx"], 1.0)
    out = str(tmp_path / "mlem")
    par = write_par(tmp_path / "mlem.par", "MLEM", **{
        "input type": "Sinogram2D", "input file": sino, "initial estimate": init,
        "output filename prefix": out, "number of iterations": g["iters"],
        "forwardprojector": "Siddon", "backprojector": "Siddon",
        "cylindrical pet radius (in mm)": g["rscanner"]})
    assert cli.main(["MLEM", par]) == 0
    assert_bitexact(read(out + "_final")[0], f["final"])
    assert_bitexact(read(out + "_sensitivity")[0], f["sens"])
    rows = log_rows(out)
    liks = rows["Likelihood por Iteración:"]
    assert len(liks) == g["iters"] + 1
    assert np.allclose(liks[:-1], f["lik"][:-1].astype(np.float64), rtol=1e-4)
    whole = np.array(rows["Tiempos de Reconstrucción por Iteración [mseg]"])
    parts = sum(np.array(rows[k]) for k in ("Tiempos de Forwardprojection por Iteración [mseg]",
                                            "Tiempos de Backwardprojection por Iteración [mseg]",
                                            "Tiempos de UpdatePixel por Iteración [mseg]"))
    assert len(whole) == g["iters"] and np.all(whole > 0)
    assert np.allclose(whole, parts, rtol=1e-4, atol=1e-3)


def _sino3d_inputs(tmp_path, g, meas):
    nsino, mn, mx = hip.segment_table(g["rings"], g["span"], g["mrd"])
    meas = np.asarray(meas, np.float32).reshape(int(nsino.sum()), g["nproj"], g["nr"])
    sino = interfile.write_sinogram3d(str(tmp_path / "sino"), meas, nsino, mn, mx, g["rings"],
                                      2.0, 180.0 / g["nproj"])
    init = interfile.write_image(str(tmp_path / "init"),
                                 np.ones((g["nz"], g["ny"], g["nx"]), np.float32),
                                 g["dx"], g["dx"], g["dz"])
    keys = {"input type": "Sinogram3D", "input file": sino, "initial estimate": init,
            "forwardprojector": "HipSiddonProjector", "backprojector": "HipSiddonProjector",
            "cylindrical pet radius (in mm)": g["rscanner"], "radius fov (in mm)": g["rfov"],
            "axial fov (in mm)": g["zfov"]}
    return meas, keys


def test_cli_osem3d_matches_reference_and_plan(torch, ctx, golden, tmp_path):
    log("case: CLI OSEM, Sinogram3D input, mid3d geometry, 7 subsets x 2 iterations")
    f = golden("mid3d_osem.npz")
    g = json.loads(str(f["geom"]))
    meas, keys = _sino3d_inputs(tmp_path, g, f["meas"])
    out = str(tmp_path / "osem")
    par = write_par(tmp_path / "osem.par", "OSEM", **keys, **{
        "output filename prefix": out, "number of iterations": g["iters"],
        "number of subsets": g["subsets"], "save estimates at iteration intervals": 1})
    assert cli.main(["OSEM", par]) == 0
    final = read(out + "_final")
    ref = f["final"]
    assert np.array_equal(np.isnan(final), np.isnan(ref)), "NaN positions differ"
    e = rel_l2(final, ref)
    log(f"CLI OSEM rel L2 vs reference {e:.2e}")
    assert e <= IMG_RL2
    assert_voxel_close(read(out + "_sensitivity_subset_0"), f["sens0"])

    p = plan3d(ctx, g)
    md = dev(torch, meas)
    init = torch.ones(p.image_shape(), dtype=torch.float32, device="cuda")
    for t in range(g["iters"]):
        img, _ = p.reconstruct(md, init, t + 1, subsets=g["subsets"], bp_mode=hip.BP_GATHER)
        assert_bitexact(read(f"{out}_iter_{t}"), host(img), nan_ok=True)
    assert_bitexact(final, host(img), nan_ok=True)
    for s in (0, g["subsets"] - 1):
        sens = p.sensitivity(subset=s, nsubsets=g["subsets"], mode=hip.BP_GATHER)
        assert_bitexact(read(f"{out}_sensitivity_subset_{s}"), host(sens))
    assert not os.path.exists(f"{out}_iter_{g['iters']}.h33")


def test_cli_osem3d_with_corrections_equals_plan(torch, ctx, tmp_path):
    log("case: CLI OSEM with multiplicative and additive sinograms, tiny 3D geometry")
    g = dict(nx=32, ny=32, nz=9, dx=4.0, dz=4.0, nproj=24, nr=40, rings=5, rfov=64.0, zfov=36.0,
             rscanner=300.0, span=3, mrd=2)
    nsino, mn, mx = hip.segment_table(g["rings"], g["span"], g["mrd"])
    rng = np.random.default_rng(11)
    shape = (int(nsino.sum()), g["nproj"], g["nr"])
    meas = rng.poisson(20.0, shape).astype(np.float32)
    mult = rng.uniform(1.0, 5.0, shape).astype(np.float32)
    add = rng.uniform(0.0, 2.0, shape).astype(np.float32)
    meas, keys = _sino3d_inputs(tmp_path, g, meas)
    interfile.write_sinogram3d(str(tmp_path / "mult"), mult, nsino, mn, mx, g["rings"], 2.0, 7.5)
    interfile.write_sinogram3d(str(tmp_path / "add"), add, nsino, mn, mx, g["rings"], 2.0, 7.5)
    out = str(tmp_path / "corr")
    par = write_par(tmp_path / "corr.par", "OSEM", **keys, **{
        "output filename prefix": out, "number of iterations": 3, "number of subsets": 4,
        "multiplicative sinogram": str(tmp_path / "mult.h33"),
        "additive sinogram": str(tmp_path / "add.h33")})
    assert cli.main(["OSEM", par]) == 0
    p = plan3d(ctx, g)
    init = torch.ones(p.image_shape(), dtype=torch.float32, device="cuda")
    img, lik = p.reconstruct(dev(torch, meas), init, 3, subsets=4, bp_mode=hip.BP_GATHER,
                             osem_rules=1, additive=dev(torch, add), mult=dev(torch, mult))
    assert_bitexact(read(out + "_final"), host(img), nan_ok=True)
    liks = log_rows(out)["Likelihood por Iteración:"]
    assert len(liks) == 3 * 4 + 1
    assert np.allclose(liks, lik.astype(np.float32), rtol=1e-5, equal_nan=True)


def test_mlem2d_multiple_equals_reference(torch, ctx, golden, tmp_path):
    """Slice-by-slice 2D MLEM, all slices in one batch: the volume equals the
    reference's Mlem2dMultiple bit for bit, through the library call and through the
    command with a multi-slice Interfile input (input type Sinograms2D)."""
    from apirl_amd import multislice
    log("case: Mlem2dMultiple, 4 slices 64x64, 10 iterations, batched")
    f = golden("mlem2d_multiple.npz")
    g = json.loads(str(f["geom"]))
    sinos = f["sinos"].astype(np.float32)
    plan = multislice.slice_plan(ctx, g["nx"], g["dx"], g["nproj"], g["nr"], g["rscanner"])
    vol = multislice.mlem2d_multiple(plan, dev(torch, sinos), g["iters"])
    torch.cuda.synchronize()
    assert_bitexact(host(vol), f["volume"])

    hdr = interfile.write_image(str(tmp_path / "sinos"), sinos, 1.0, 1.0, 1.0)
    init = interfile.write_image(str(tmp_path / "init"),
                                 np.zeros((g["nslices"], g["nx"], g["nx"]), np.float32),
                                 g["dx"], g["dx"], 4.0)
    out = str(tmp_path / "ms")
    par = write_par(tmp_path / "ms.par", "MLEM", **{
        "input type": "Sinograms2D", "input file": hdr, "initial estimate": init,
        "output filename prefix": out, "number of iterations": g["iters"],
        "save estimates at iteration intervals": 5,
        "forwardprojector": "Siddon", "backprojector": "Siddon",
        "cylindrical pet radius (in mm)": g["rscanner"], "radius fov (in mm)": g["rfov"],
        "axial fov (in mm)": g["zfov"]})
    assert cli.main(["MLEM", par]) == 0
    assert_bitexact(read(out + "_volFinal"), f["volume"])
    for k in range(g["nslices"]):
        assert_bitexact(read(f"{out}_slice_{k}_final")[0], f["volume"][k])
    # iteration 5 (0-based) was saved: it equals a 6-iteration run
    vol6 = multislice.mlem2d_multiple(plan, dev(torch, sinos), 6)
    torch.cuda.synchronize()
    assert_bitexact(read(f"{out}_slice_2_iter_5")[0], host(vol6)[2])


def test_cli_oversampled_osem_equals_plan(torch, ctx, tmp_path):
    """.par keys "siddon number of (axial) samples on the detector" trace every bin as
    2 x 2 LORs; the command's OSEM (gather backprojection) equals Plan.reconstruct
    with the same oversampling bit for bit."""
    log("case: CLI OSEM with 2 x 2 LORs per bin, tiny 3D geometry")
    g = dict(nx=32, ny=32, nz=9, dx=4.0, dz=4.0, nproj=24, nr=40, rings=5, rfov=64.0, zfov=36.0,
             rscanner=300.0, span=3, mrd=2)
    nsino, _, _ = hip.segment_table(g["rings"], g["span"], g["mrd"])
    rng = np.random.default_rng(12)
    meas = rng.poisson(20.0, (int(nsino.sum()), g["nproj"], g["nr"])).astype(np.float32)
    meas, keys = _sino3d_inputs(tmp_path, g, meas)
    out = str(tmp_path / "os")
    par = write_par(tmp_path / "os.par", "OSEM", **keys, **{
        "output filename prefix": out, "number of iterations": 2, "number of subsets": 4,
        "siddon number of samples on the detector": 2,
        "siddon number of axial samples on the detector": 2})
    assert cli.main(["OSEM", par]) == 0
    p = plan3d(ctx, g)
    p.set_oversampling(2, 2)
    init = torch.ones(p.image_shape(), dtype=torch.float32, device="cuda")
    img, _ = p.reconstruct(dev(torch, meas), init, 2, subsets=4, bp_mode=hip.BP_GATHER,
                           osem_rules=1)
    got, ref = read(out + "_final"), host(img)
    assert np.array_equal(got.view(np.int32), ref.view(np.int32))


def test_cli_mmr_mlem_equals_plan(torch, ctx, golden, tmp_path):
    """Input type Sinogram3DSiemensMmr: the command builds the mMR geometry from the
    scanner's constants and the header's segment table, and its reconstruction equals
    Plan.reconstruct on the same mMR plan within the image bar.  MLEM: with 6 angles
    per subset the coarse grid's corner voxels see no LOR of a subset, and OSEM's
    threshold-0 update then spreads 0/0 = NaN through the image, as in the reference."""
    log("case: CLI MLEM, Sinogram3DSiemensMmr input")
    f = golden("mmr.npz")
    g = json.loads(str(f["geom"]))
    nsino, mn, mx = hip.segment_table(g["rings"], g["span"], g["mrd"])
    meas = f["meas"].astype(np.float32).reshape(int(nsino.sum()), g["nproj"], g["nr"])
    sino = interfile.write_sinogram3d(str(tmp_path / "mmr"), meas, nsino, mn, mx, g["rings"],
                                      8.18, 180.0 / g["nproj"])
    init = interfile.write_image(str(tmp_path / "init"),
                                 np.ones((g["nz"], g["ny"], g["nx"]), np.float32),
                                 g["dx"], g["dx"], g["dz"])
    out = str(tmp_path / "mmr_out")
    par = write_par(tmp_path / "mmr.par", "MLEM", **{
        "input type": "Sinogram3DSiemensMmr", "input file": sino, "initial estimate": init,
        "output filename prefix": out, "number of iterations": 3,
        "forwardprojector": "Siddon", "backprojector": "Siddon"})
    assert cli.main(["MLEM", par]) == 0
    p = hip.Plan.for_3d(ctx, g["nx"], g["ny"], g["nz"], g["dx"], g["dz"], g["nproj"], g["nr"],
                        297.0, 328.0, g["rings"], 260.0, g["span"], g["mrd"])
    p.set_scanner("mmr")
    x0 = torch.ones(p.image_shape(), dtype=torch.float32, device="cuda")
    img, _ = p.reconstruct(dev(torch, meas), x0, 3, subsets=1, bp_mode=hip.BP_GATHER,
                           osem_rules=0)
    got, ref = read(out + "_final"), host(img)
    assert np.isfinite(got).all() and np.isfinite(ref).all() and ref.max() > 0
    assert rel_l2(got, ref) <= IMG_RL2


def test_cli_osem_sensitivity_filename_reload(torch, ctx, golden, tmp_path):
    """"sensitivity filename" (OsemSinogram3d.cpp:83-95): a second OSEM run that reads
    the first run's <prefix>_sensitivity_subset_<s>.h33 gives the same images bit for
    bit and writes no sensitivity of its own."""
    log("case: CLI OSEM, sensitivity images reloaded from the first run")
    f = golden("mid3d_osem.npz")
    g = json.loads(str(f["geom"]))
    meas, keys = _sino3d_inputs(tmp_path, g, f["meas"])
    common = dict(keys, **{"number of iterations": g["iters"], "number of subsets": g["subsets"]})
    a, b = str(tmp_path / "first"), str(tmp_path / "second")
    assert cli.main(["OSEM", write_par(tmp_path / "a.par", "OSEM", **common,
                                       **{"output filename prefix": a})]) == 0
    assert cli.main(["OSEM", write_par(tmp_path / "b.par", "OSEM", **common,
                                       **{"output filename prefix": b,
                                          "sensitivity filename": a + "_sensitivity"})]) == 0
    assert_bitexact(read(b + "_final"), read(a + "_final"), nan_ok=True)
    assert not os.path.exists(b + "_sensitivity_subset_0.h33")


def test_cli_project_backproject_vs_reference(torch, ctx, golden, tmp_path):
    """project / backproject (cmd/src/project.cpp, backproject.cpp) on a Sinogram3D:
    the whole projection bit-exact and the backprojection within 1e-5 per voxel of
    the reference's SiddonProjector (mid3d goldens); subset 2 of 7 bit-exact against
    the oracle's projection of the subset's angles."""
    import _oracle as O
    log("case: CLI project / backproject, Sinogram3D, mid3d geometry")
    f = golden("mid3d_project_backproject.npz")
    g = json.loads(str(f["geom"]))
    meas, keys = _sino3d_inputs(tmp_path, g, f["meas"])
    img = interfile.write_image(str(tmp_path / "img"), f["img"].reshape(g["nz"], g["ny"], g["nx"]),
                                g["dx"], g["dx"], g["dz"])
    cyl = {k: keys[k] for k in ("cylindrical pet radius (in mm)", "radius fov (in mm)",
                                "axial fov (in mm)")}
    out = str(tmp_path / "proj")
    assert cli.main(["project", write_par(tmp_path / "p.par", "Projection", **cyl, **{
        "input file": img, "output type": "Sinogram3D", "projector": "Siddon",
        "output filename": out, "output projection": keys["input file"]})]) == 0
    got = interfile.read_sinogram3d(out + ".h33")["bins"]
    assert_bitexact(got.ravel(), f["proj"])
    out2 = str(tmp_path / "proj_s2")
    assert cli.main(["project", write_par(tmp_path / "p2.par", "Projection", **cyl, **{
        "input file": img, "output type": "Sinogram3D", "projector": "CuSiddonProjector",
        "output filename": out2, "output projection": keys["input file"],
        "number of subsets": 7, "subset index": 2})]) == 0
    sub = interfile.read_sinogram3d(out2 + ".h33")["bins"]
    im = O.image_desc(g["nx"], g["ny"], g["nz"], g["dx"], g["dz"])
    ang, r = O.sino_grid(g["nproj"], g["nr"], g["rfov"])
    za, zb = O.ring_table(g["rings"], g["zfov"], g["span"], g["mrd"])
    idx = O.subset_angles(g["nproj"], 7, 2)
    ref = O.project3d(im, f["img"], np.ascontiguousarray(ang[idx]), r, g["rscanner"], za, zb)
    assert_bitexact(sub.ravel(), ref.ravel())
    bimg = interfile.write_image(str(tmp_path / "zero"), np.zeros((g["nz"], g["ny"], g["nx"]),
                                 np.float32), g["dx"], g["dx"], g["dz"])
    bout = str(tmp_path / "back")
    assert cli.main(["backproject", write_par(tmp_path / "b.par", "Backproject", **cyl, **{
        "input type": "Sinogram3D", "input file": keys["input file"],
        "backprojector": "Siddon", "output image": bimg, "output filename": bout})]) == 0
    assert_voxel_close(read(bout), f["back"])


def test_cli_mmr_subset_projection_and_cumlem(torch, ctx, golden, tmp_path):
    """project onto subset 1 of 3 of an mMR sinogram: bit-exact against the
    reference's projection of that getSubset copy.  cuMLEM (cmdGPU/src/cuMLEM.cpp) with
    its GPU keys on Sinogram3DSiemensMmr input: MLEM without "number of subsets"
    (the reference's MlemSinogram3d golden), OSEM with it (its OsemSinogram3d golden),
    both within the image bar."""
    log("case: CLI mMR subset projection, cuMLEM MLEM and OSEM")
    c = golden("corrections.npz")
    gm = json.loads(str(c["geom_mmr"]))
    mm = golden("mmr.npz")
    nsino, mn, mx = hip.segment_table(gm["rings"], gm["span"], gm["mrd"])
    meas = mm["meas"].astype(np.float32).reshape(int(nsino.sum()), gm["nproj"], gm["nr"])
    sino = interfile.write_sinogram3d(str(tmp_path / "mmr"), meas, nsino, mn, mx, gm["rings"],
                                      8.18, 180.0 / gm["nproj"])
    shp = (gm["nz"], gm["ny"], gm["nx"])
    img = interfile.write_image(str(tmp_path / "img"), mm["img"].reshape(shp), gm["dx"], gm["dx"],
                                gm["dz"])
    out = str(tmp_path / "sub")
    assert cli.main(["project", write_par(tmp_path / "p.par", "Projection", **{
        "input file": img, "output type": "Sinogram3DSiemensMmr", "projector": "Siddon",
        "output filename": out, "output projection": sino, "number of subsets": 3,
        "subset index": 1})]) == 0
    got = interfile.read_sinogram3d(out + ".h33")["bins"]
    assert_bitexact(got.ravel(), c["mmr_subset1of3_proj"])
    init = interfile.write_image(str(tmp_path / "init"), np.ones(shp, np.float32), gm["dx"],
                                 gm["dx"], gm["dz"])
    gpu = {"projector block size": "{128,1,1}", "backprojector block size": "{128, 1, 1}",
           "pixel update block size": "{512,1,1}", "gpu id": 0,
           "forwardprojector": "CuSiddonProjector", "backprojector": "CuSiddonProjector",
           "input type": "Sinogram3DSiemensMmr", "input file": sino, "initial estimate": init}
    m = golden("mmr_mlem3d.npz")
    out = str(tmp_path / "cumlem")
    assert cli.main(["cuMLEM", write_par(tmp_path / "m.par", "MLEM", **gpu, **{
        "output filename prefix": out, "number of iterations": 3})]) == 0
    assert rel_l2(read(out + "_final"), m["final"]) <= IMG_RL2
    # OSEM: the mMR OSEM case (192 mm image inside the radial FOV, seeded counts)
    go = json.loads(str(c["geom_mmr_osem"]))
    ns2, mn2, mx2 = hip.segment_table(go["rings"], go["span"], go["mrd"])
    counts = np.random.default_rng(go["seed"]).poisson(
        2.0, (int(ns2.sum()), go["nproj"], go["nr"])).astype(np.float32)
    osino = interfile.write_sinogram3d(str(tmp_path / "mmro"), counts, ns2, mn2, mx2, go["rings"],
                                       4.09, 180.0 / go["nproj"])
    oshp = (go["nz"], go["ny"], go["nx"])
    oinit = interfile.write_image(str(tmp_path / "oinit"), np.ones(oshp, np.float32), go["dx"],
                                  go["dx"], go["dz"])
    out = str(tmp_path / "cuosem")
    assert cli.main(["cuMLEM", write_par(tmp_path / "o.par", "MLEM", **dict(
        gpu, **{"input file": osino, "initial estimate": oinit}), **{
        "output filename prefix": out, "number of iterations": go["iters"],
        "number of subsets": go["subsets"]})]) == 0
    final, ref = read(out + "_final"), c["mmr_osem_final"]
    assert np.array_equal(np.isnan(final), np.isnan(ref))
    assert rel_l2(final, ref) <= IMG_RL2
    # a block size without its braces is refused, as the reference refuses it
    bad = dict(gpu, **{"projector block size": "128,1,1"})
    assert cli.main(["cuMLEM", write_par(tmp_path / "bad.par", "MLEM", **bad, **{
        "output filename prefix": out, "number of iterations": 1})]) == 1


def test_cli_project_2d_subset_vs_reference(torch, ctx, golden, tmp_path):
    """project / backproject on Sinogram2D (the 2D LOR form with the r <= rfov
    rule): the whole projection bit-exact against the reference's (c1 golden) and
    the gather backprojection bit-exact as well."""
    log("case: CLI project / backproject, Sinogram2D, c1 geometry")
    f = golden("c1_project_backproject.npz")
    g = json.loads(str(f["geom"]))
    img = interfile.write_image(str(tmp_path / "img"),
                                f["img"].reshape(1, g["nx"], g["nx"]), g["dx"], g["dx"], 1.0)
    nproj, nr = g["nproj"], g["nr"]
    sample = interfile.write_sinogram2d(str(tmp_path / "s"), np.zeros((nproj, nr), np.float32))
    cyl = {"radius fov (in mm)": g["rfov"], "cylindrical pet radius (in mm)": g["rscanner"],
           "axial fov (in mm)": 0}
    out = str(tmp_path / "p")
    assert cli.main(["project", write_par(tmp_path / "p.par", "Projection", **cyl, **{
        "input file": img, "output type": "Sinogram2D", "projector": "Siddon",
        "output filename": out, "output projection": sample})]) == 0
    assert_bitexact(interfile.read_sinogram2d(out + ".h33").ravel(), f["proj"])
