"""Command-line reconstruction from the reference's .par files, on the GPU path.

    python -m apirl_amd.cli OSEM recon.par
    python -m apirl_amd.cli MLEM recon.par

The commands mirror cmd/src/OSEM.cpp and cmd/src/MLEM.cpp for the inputs this path
covers.
  * Input types: "Sinogram3D" (cylindrical scanner), "Sinogram3DSiemensMmr" (the mMR's
    LOR geometry and constants, scatter backprojection), "Sinogram2D" (2D MLEM under
    either command, as in the reference) and "Sinograms2D" /
    "Sinograms2DinCylindrical3Dpet" (slice-by-slice 2D MLEM, Mlem2dMultiple, run as
    one batch; outputs <prefix>_slice_<i>_{sensitivity,final,iter_<t>}.h33 and
    <prefix>_volFinal.h33).
  * Keys, all read with the reference's parser rules (apirl_amd.interfile):
      "input type", "input file", "initial estimate", "output filename prefix",
      "number of iterations", "number of subsets" (OSEM);
      "cylindrical pet radius (in mm)", "radius fov (in mm)" and "axial fov (in mm)"
      (data/src/readParameters.cpp:281-330);
      "forwardprojector" and "backprojector" (required, as in the reference);
      optional "save estimates at iteration intervals", "multiplicative sinogram",
      "additive sinogram".  "enforce initial positivity condition" is accepted and,
      as in the reference, has no effect (readParameters.cpp:242-270).
  * Projectors: "Siddon" and "HipSiddonProjector" select this path; any other name
    is refused, and so are "sensitivity filename" and "attenuation image filename".
    "siddon number of samples on the detector" / "... axial samples ..." trace 3D
    bins as several LORs (apirl_plan_set_oversampling), backprojected with the
    scatter form; the backprojector is otherwise the deterministic gather form.
  * Outputs, named as in OsemSinogram3d.cpp:110-113, 270-286:
      <prefix>_sensitivity_subset_<s>.h33 (MLEM: <prefix>_sensitivity.h33),
      <prefix>_iter_<t>.h33 at the save interval, <prefix>_final.h33, and
      <prefix>.log in the reference's log format (parameters, saved images, total
      time, per sub-iteration times and likelihoods).
The arithmetic is the reference's (APIRL_PREC_REFERENCE)."""
import os
import re
import sys
import time

import numpy as np

from . import hip, multislice
from .interfile import (FormatError, ParFile, atof, atoi, read_image, read_sinogram2d,
                        read_sinogram3d, write_image, write_sinogram2d, write_sinogram3d)

PROJECTORS = ("Siddon", "HipSiddonProjector")


def _load_sino(par, key, kind3d, torch):
    name = par.get(key)
    if not name:
        return None
    if kind3d:
        s = read_sinogram3d(name)["bins"]
    else:
        s = read_sinogram2d(name)
    return torch.from_numpy(np.ascontiguousarray(s)).cuda()


def _sens_files(par, is3d, osem, subsets):
    """The reference's "sensitivity filename" (readParameters.cpp:25-54): OSEM reads
    <name>_subset_<s>.h33 per subset (OsemSinogram3d.cpp:83-95), the MLEM drivers the
    named image itself (MlemSinogram3d.cpp:132-138, Mlem2d.cpp:96-102).  None when the
    key is absent."""
    name = par.get("sensitivity filename")
    if name is None:
        return None
    paths = ([f"{name}_subset_{s}.h33" for s in range(subsets)] if (is3d and osem)
             else [name])
    return np.stack([read_image(q)[0] for q in paths]).astype(np.float32)


def run(method, par_path, log=print, gpu=0, mmr_osem_if_subsets=False, projectors=None):
    import torch

    if not par_path.endswith(".par"):
        raise FormatError("the parameter file must have the .par extension")
    par = ParFile(par_path, method)
    itype = par.get("input type", required=True)
    infile = par.get("input file", required=True)
    init_name = par.get("initial estimate", required=True)
    prefix = par.get("output filename prefix", required=True)
    iters = atoi(par.get("number of iterations", required=True))
    subsets = atoi(par.get("number of subsets", required=True)) if method == "OSEM" else 1
    if mmr_osem_if_subsets and itype == "Sinogram3DSiemensMmr":
        # cuMLEM.cpp:415-420: OSEM on mMR input when "number of subsets" is given
        n = par.get("number of subsets")
        if n is not None and atoi(n) > 0:
            method, subsets = "OSEM", atoi(n)
    interval = atoi(par.get("save estimates at iteration intervals", "0"))
    if iters < 0 or subsets < 1 or interval < 0:
        raise FormatError(f"{par_path}: negative iterations or interval, or no subsets")
    # both projector names are required (readParameters.cpp:117-158)
    projectors = PROJECTORS if projectors is None else projectors
    for key in ("forwardprojector", "backprojector"):
        name = par.get(key, required=True).strip()
        if name not in projectors:
            raise FormatError(f"{key} '{name}' is not available on this path "
                              f"(use one of {', '.join(projectors)})")
    # LORs per bin (readParameters.cpp:162-203; absent keys mean one)
    samples = []
    for key in ("siddon number of samples on the detector",
                "siddon number of axial samples on the detector"):
        v = par.get(key)
        n = 1 if v is None else atoi(v)
        if n < 1:
            raise FormatError(f"'{key}' = {v.strip()}: at least one sample per bin")
        samples.append(n)
    if par.get("attenuation image filename") is not None:
        raise FormatError("'attenuation image filename' is not available on this path: the "
                          "attenuation enters as the multiplicative sinogram")
    init, (dx, dy, dz) = read_image(init_name)
    nz, ny, nx = init.shape
    ctx = hip.Context(gpu)
    if itype == "Sinogram3D":
        rscanner = atof(par.get("cylindrical pet radius (in mm)", required=True))
        rfov = atof(par.get("radius fov (in mm)", required=True))
        zfov = atof(par.get("axial fov (in mm)", required=True))
        sino = read_sinogram3d(infile)
        z1, z2 = hip.ring_table_from(sino["rings"], zfov, sino["nsino"], sino["minrd"],
                                     sino["maxrd"])
        plan = hip.Plan(ctx, nx, ny, nz, dx, dz, sino["nproj"], sino["nr"], rfov, rscanner,
                        dims=3, z1=z1, z2=z2,
                        ring_width=np.float32(zfov) / np.float32(sino["rings"]))
        if samples != [1, 1]:
            plan.set_oversampling(*samples)
        meas = torch.from_numpy(sino["bins"]).cuda()
        is3d = True
    elif itype == "Sinogram3DSiemensMmr":
        # OSEM.cpp:395-403: the scanner's constants, no cylindrical keys
        # (Sinogram3DSiemensMmr.cpp:9-27); the mMR LOR geometry (apirl_plan_set_scanner)
        sino = read_sinogram3d(infile)
        zfov = 260.0
        z1, z2 = hip.ring_table_from(sino["rings"], zfov, sino["nsino"], sino["minrd"],
                                     sino["maxrd"])
        plan = hip.Plan(ctx, nx, ny, nz, dx, dz, sino["nproj"], sino["nr"], 297.0, 328.0,
                        dims=3, z1=z1, z2=z2,
                        ring_width=np.float32(zfov) / np.float32(sino["rings"]))
        plan.set_scanner("mmr")
        if samples != [1, 1]:
            raise FormatError("oversampled LORs are defined for the cylindrical scanner")
        meas = torch.from_numpy(sino["bins"]).cuda()
        is3d = True
    elif itype == "Sinogram2D":
        # Mlem2d takes the radial FOV from the image, (dx * nx) / 2 in float
        # (Mlem2d.cpp:12, Images.cpp:378); the reference leaves the scanner radius of
        # this input type unset, here it is "cylindrical pet radius (in mm)" or 400 mm
        rscanner = atof(par.get("cylindrical pet radius (in mm)", "400"))
        s2 = read_sinogram2d(infile)
        nproj, nr = s2.shape
        rfov = float(np.float32(np.float32(dx) * np.float32(nx)) / np.float32(2))
        plan = hip.Plan.for_2d(ctx, nx, dx, nproj, nr, rfov, rscanner)
        meas = torch.from_numpy(np.ascontiguousarray(s2)).cuda()
        is3d = False
        # both commands run Mlem2d on this input: no subsets, MLEM rules
        # (OSEM.cpp:360-369, MLEM.cpp:356-366)
        subsets = 1
    elif itype in ("Sinograms2D", "Sinograms2DinCylindrical3Dpet"):
        # MLEM.cpp:429-440 / OSEM.cpp:384-392: Mlem2dMultiple over the slices
        for key in ("cylindrical pet radius (in mm)", "radius fov (in mm)", "axial fov (in mm)"):
            par.get(key, required=True)
        rscanner = atof(par.get("cylindrical pet radius (in mm)"))
        sinos, _ = read_image(infile)
        if sinos.shape[0] != nz:
            raise FormatError(f"{infile}: {sinos.shape[0]} 2D sinograms for an initial "
                              f"estimate of {nz} slices")
        if par.get("sensitivity filename") is not None:
            raise FormatError("'sensitivity filename' is not available for slice-by-slice "
                              "input")
        plan = multislice.slice_plan(ctx, nx, dx, sinos.shape[1], sinos.shape[2], rscanner)
        return _run_multislice(par, method, plan, sinos, iters, interval, prefix, infile,
                               (dx, dy, dz), torch, log)
    else:
        raise FormatError(f"input type '{itype}' is not available on this path")
    mult = _load_sino(par, "multiplicative sinogram", is3d, torch)
    add = _load_sino(par, "additive sinogram", is3d, torch)
    x = torch.from_numpy(np.ascontiguousarray(init)).cuda()
    osem_rules = 1 if method == "OSEM" and is3d else 0
    if add is not None and osem_rules and subsets > 1:
        # OsemSinogram3d.cpp:172-173 adds the full additive sinogram to each subset
        # sinogram with Sinogram3D::addBinToBin (Sinogram3D.cpp:857-873), which reads
        # the additive bin at the subset's own angle index: the reference's result is
        # reproduced, and said so
        print("note: as the reference's OSEM, subset angle k takes the additive "
              "sinogram's angle k (not the subset's own angle s + S k)", file=sys.stderr)
    # the deterministic gather backprojector, oversampled 3D bins included (their LORs
    # gathered in turn); 2D projectors have no oversampling (SiddonProjector.cpp:
    # 34-193), so there the keys have no effect
    bp_mode = hip.BP_GATHER
    sens = torch.empty((subsets,) + plan.image_shape(), dtype=torch.float32, device="cuda")
    given = _sens_files(par, is3d, osem_rules, subsets)
    if given is not None:
        if given.size != sens.numel():
            raise FormatError("the sensitivity images do not match the initial estimate")
        sens.copy_(torch.from_numpy(given.reshape(sens.shape)))
    t0 = time.perf_counter()
    liks, times, saved = [], [], []
    done = 0
    # iteration t (0-based) is saved when t % interval == 0 (OsemSinogram3d.cpp:266-276):
    # run up to each save point, save, continue; the state between chunks is the image
    while True:
        step = iters - done
        if interval:
            nxt = 1 if done == 0 else ((done - 1) // interval + 1) * interval + 1
            step = min(step, nxt - done)
        tm = np.zeros((4, step * subsets), np.float32)
        img, lik = plan.reconstruct(meas, x, step, subsets=subsets, bp_mode=bp_mode,
                                    osem_rules=osem_rules, additive=add, mult=mult,
                                    sens_out=None if given is not None else sens,
                                    sens_in=sens if given is not None else None,
                                    times_out=tm if step else None)
        liks.extend(lik[:step * subsets].tolist())
        times.append(tm)
        final_lik = float(lik[step * subsets])
        x = img
        done += step
        if step and interval and (done - 1) % interval == 0:
            write_image(f"{prefix}_iter_{done - 1}", img.cpu().numpy(), dx, dy, dz)
            saved.append(done - 1)
        if done >= iters:
            break
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    s_host = sens.cpu().numpy()
    if given is not None:
        pass  # images read from files are not written again (OsemSinogram3d.cpp:83-95)
    elif method == "OSEM" and is3d:
        for s in range(subsets):
            write_image(f"{prefix}_sensitivity_subset_{s}", s_host[s], dx, dy, dz)
    else:
        write_image(f"{prefix}_sensitivity", s_host[0], dx, dy, dz)
    final = write_image(f"{prefix}_final", x.cpu().numpy(), dx, dy, dz)
    head = [("Iterations", str(iters)),
            ("Input Projections", "3D Sinogram" if is3d else "2D Sinogram"),
            ("Size of Sinogram2D", f"{meas.shape[-1]}[r] x {meas.shape[-2]}[ang]")]
    if is3d:
        head += [("Rings", str(sino["rings"])), ("Segments", str(len(sino["nsino"])))]
    head += [("Image Dimensions", "3" if nz > 1 else "2"),
             ("Image Size", f"{nx}[x] x {ny}[y] x {nz}[z]")]
    write_log(prefix, head, saved, elapsed * 1e3, np.concatenate(times, axis=1),
              liks + [final_lik])
    log(f"{method}: {iters} iteration(s) x {subsets} subset(s) in {elapsed:.3f} s -> {final}")
    return final


def _row(values):
    """Logger::writeRowOfNumbers: floats as an ostream prints them (6 significant
    digits), separated by ", " (data/src/Logger.cpp:119-135)."""
    return ", ".join(format(float(np.float32(v)), "g") for v in values)


def write_log(prefix, head, saved, total_ms, times, liks):
    """<prefix>.log in the reference drivers' format (OsemSinogram3d.cpp:125-150,
    266-317; Logger.cpp: "label: value" lines, CR LF line ends): the parameters, the
    saved images, the total time, per sub-iteration times in ms (whole, projection,
    backprojection, update; measured with HIP events here) and the likelihoods."""
    lines = ["######## ML-EM Reconstruction #########", f"Name: {prefix}", "Type: ML-EM"]
    lines += [f"{k}: {v}" for k, v in head]
    lines.append("_______RECONSTRUCCION_______")
    lines += [f"Imagen de iteración {t} guardada en: {prefix}_iter_{t}" for t in saved]
    lines.append(f"Imagen final guardada en: {prefix}_final")
    lines.append("_______RESULTADOS DE RECONSTRUCCION_______")
    lines.append(f"Tiempo Total de Reconstrucción:: {total_ms:f}")
    for title, row in zip(("Tiempos de Reconstrucción por Iteración [mseg]",
                           "Tiempos de Forwardprojection por Iteración [mseg]",
                           "Tiempos de Backwardprojection por Iteración [mseg]",
                           "Tiempos de UpdatePixel por Iteración [mseg]"), times):
        lines += [title, _row(row)]
    lines += ["Likelihood por Iteración:", _row(liks)]
    with open(prefix + ".log", "wb") as f:
        f.write("".join(line + "\r\n" for line in lines).encode("utf-8"))


def _run_multislice(par, method, plan, sinos, iters, interval, prefix, infile, vox, torch,
                    log):
    """Mlem2dMultiple (Mlem2dMultiple.cpp:64-84): each slice's Mlem2d writes its own
    files under <prefix>_slice_<i>, and the stacked volume goes to <prefix>_volFinal."""
    dx, dy, dz = vox
    nz = sinos.shape[0]

    def load(key):
        name = par.get(key)
        if not name:
            return None
        s, _ = read_image(name)
        if s.shape != sinos.shape:
            raise FormatError(f"{name}: shape {s.shape}, the sinograms are {sinos.shape}")
        return torch.from_numpy(np.ascontiguousarray(s)).cuda()

    mult = load("multiplicative sinogram")
    add = load("additive sinogram")
    meas = torch.from_numpy(np.ascontiguousarray(sinos)).cuda()
    t0 = time.perf_counter()
    vol, done = None, 0
    while True:
        step = iters - done
        if interval:
            nxt = 1 if done == 0 else ((done - 1) // interval + 1) * interval + 1
            step = min(step, nxt - done)
        vol = multislice.mlem2d_multiple(plan, meas, step, init=vol, mult=mult, additive=add)
        done += step
        if step and interval and (done - 1) % interval == 0:
            host = vol.cpu().numpy()
            for k in range(nz):
                write_image(f"{prefix}_slice_{k}_iter_{done - 1}", host[k], dx, dy, dz)
        if done >= iters:
            break
    sens = multislice.sensitivities(plan, nz, mult=mult)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    host, s_host = vol.cpu().numpy(), sens.cpu().numpy()
    for k in range(nz):
        write_image(f"{prefix}_slice_{k}_sensitivity", s_host[k], dx, dy, dz)
        write_image(f"{prefix}_slice_{k}_final", host[k], dx, dy, dz)
    final = write_image(f"{prefix}_volFinal", host, dx, dy, dz)
    with open(prefix + ".log", "w") as f:
        f.write(f"Name := {prefix}\n")
        f.write(f"Method := {method} (slice by slice)\n")
        f.write(f"Iterations := {iters}\n")
        f.write(f"Slices := {nz}\n")
        f.write(f"Input := {infile}\n")
        f.write(f"Reconstruction time (s) := {elapsed:.6f}\n")
        f.write(f"Final image := {final}\n")
    log(f"{method}: {nz} slices x {iters} iteration(s) in {elapsed:.3f} s -> {final}")
    return final


# ---- project / backproject (cmd/src/project.cpp, cmd/src/backproject.cpp) ----------

def _check_par(par_path):
    if not par_path.endswith(".par"):
        raise FormatError("the parameter file must have the .par extension")


def _subset_keys(par):
    """"number of subsets" and "subset index", read together as the reference reads
    them with parametersFile_readMultipleKeys (project.cpp:226-237): both present
    give (S, index), otherwise the whole sinogram (0, 0)."""
    n, k = par.read("number of subsets"), par.read("subset index")
    if n[0] == 0 and k[0] == 0:
        S, idx = atoi(n[1]), atoi(k[1])
        if S < 0 or (S > 0 and not 0 <= idx < S):
            raise FormatError(f"subset index {idx} of {S} subsets")
        return S, idx
    return 0, 0


def _siddon_samples(par, name):
    """getSiddonProjectorParameters (readParameters.cpp:162-203): absent keys mean one
    LOR per bin."""
    out = []
    for key in ("siddon number of samples on the detector",
                "siddon number of axial samples on the detector"):
        v = par.get(key)
        n = 1 if v is None else atoi(v)
        if n < 1:
            raise FormatError(f"'{key}' = {v.strip()}: at least one sample per bin")
        out.append(n)
    return out


GPU_PROJECTORS = PROJECTORS + ("CuSiddonProjector",)


def _projector(par, key):
    name = par.get(key, required=True).strip()
    if name not in GPU_PROJECTORS:
        raise FormatError(f"{key} '{name}' is not available on this path "
                          f"(use one of {', '.join(GPU_PROJECTORS)})")
    return name


def _gpu_id(par):
    """"gpu id" (readCudaParameters.cpp:148-175): the device index, 0 if absent."""
    v = par.get("gpu id")
    return 0 if v is None else atoi(v)


def _cyl(par):
    """getCylindricalScannerParameters (readParameters.cpp:281-330)."""
    return (atof(par.get("radius fov (in mm)", required=True)),
            atof(par.get("axial fov (in mm)", required=True)),
            atof(par.get("cylindrical pet radius (in mm)", required=True)))


def _plan3d(ctx, itype, sino, par, img_shape, vox, samples):
    """Plan of a 3D sinogram of input / output type `itype` (Sinogram3D on the
    cylindrical keys, Sinogram3DSiemensMmr on the scanner's constants)."""
    nz, ny, nx = img_shape
    dx, _, dz = vox
    if itype == "Sinogram3D":
        rfov, zfov, rscanner = _cyl(par)
    else:
        rfov, zfov, rscanner = 297.0, 260.0, 328.0
    z1, z2 = hip.ring_table_from(sino["rings"], zfov, sino["nsino"], sino["minrd"],
                                 sino["maxrd"])
    plan = hip.Plan(ctx, nx, ny, nz, dx, dz, sino["nproj"], sino["nr"], rfov, rscanner, dims=3,
                    z1=z1, z2=z2, ring_width=np.float32(zfov) / np.float32(sino["rings"]))
    if itype == "Sinogram3DSiemensMmr":
        plan.set_scanner("mmr")
        if samples != [1, 1]:
            raise FormatError("oversampled LORs are defined for the cylindrical scanner")
    elif samples != [1, 1]:
        plan.set_oversampling(*samples)
    return plan


def run_project(par_path, log=print):
    """project.cpp:190-467: projects the image of "input file" onto a sinogram shaped
    as "output projection" (of "output type"), or onto subset "subset index" of
    "number of subsets" of it, and writes "output filename"."""
    import torch
    _check_par(par_path)
    par = ParFile(par_path, "Projection")
    image_name = par.get("input file", required=True)
    otype = par.get("output type", required=True)
    _projector(par, "projector")
    out_name = par.get("output filename", required=True)
    sample = par.get("output projection", required=True)
    S, sidx = _subset_keys(par)
    samples = _siddon_samples(par, "projector")
    img, (dx, dy, dz) = read_image(image_name)
    nz, ny, nx = img.shape
    ctx = hip.Context(_gpu_id(par))
    t0 = time.perf_counter()
    if otype == "Sinogram2D":
        # project.cpp:294-313: Sinogram2DinCylindrical3Dpet(sample, rfov, rscanner),
        # or its subset constructor
        rfov = atof(par.get("radius fov (in mm)", required=True))
        rscanner = atof(par.get("cylindrical pet radius (in mm)", required=True))
        nproj, nr = read_sinogram2d(sample).shape
        if nz != 1 or nx != ny:
            raise FormatError("Sinogram2D output needs a square one-slice image")
        plan = hip.Plan.for_2d(ctx, nx, dx, nproj, nr, rfov, rscanner)
        out = plan.project(torch.from_numpy(img).cuda(), subset=sidx, nsubsets=max(S, 1))
        res = write_sinogram2d(out_name, out.cpu().numpy()[0])
    elif otype in ("Sinogram3D", "Sinogram3DSiemensMmr"):
        # project.cpp:412-462: the sample's segment table, getSubset for a subset;
        # an mMR subset carries the LORs its getSubset copy traces
        sino = read_sinogram3d(sample)
        plan = _plan3d(ctx, otype, sino, par, img.shape, (dx, dy, dz), samples)
        if otype == "Sinogram3DSiemensMmr" and S > 0:
            plan.set_scanner("mmr_subset")
        out = plan.project(torch.from_numpy(img).cuda(), subset=sidx, nsubsets=max(S, 1))
        res = write_sinogram3d(out_name, out.cpu().numpy(), sino["nsino"], sino["minrd"],
                               sino["maxrd"], sino["rings"], 0.0, 0.0)
    elif otype in ("Sinograms2D", "Sinograms2DinCylindrical3Dpet"):
        # project.cpp:362-410: every slice onto its own 2D sinogram
        rfov, _, rscanner = _cyl(par)
        sinos, _ = read_image(sample)
        if sinos.shape[0] != nz:
            raise FormatError("projection of an image into a sinograms2D with a different "
                              "number of sinograms than slices")
        plan = multislice.slice_plan(ctx, nx, dx, sinos.shape[1], sinos.shape[2], rscanner,
                                     rfov=rfov)
        out = np.stack([plan.project(torch.from_numpy(np.ascontiguousarray(img[k:k + 1])).cuda())
                        .cpu().numpy()[0] for k in range(nz)])
        res = write_image(out_name, out, 1.0, 1.0, 1.0)
    else:
        raise FormatError(f"output type '{otype}' is not available on this path")
    log(f"Projection finished. Processing time: {time.perf_counter() - t0}sec.")
    return res


def run_backproject(par_path, log=print):
    """backproject.cpp:214-735: backprojects "input file" (of "input type"), or its
    subset, into an image shaped as "output image" and writes "output filename"."""
    import torch
    _check_par(par_path)
    par = ParFile(par_path, "Backproject")
    itype = par.get("input type", required=True)
    infile = par.get("input file", required=True)
    _projector(par, "backprojector")
    out_image = par.get("output image", required=True)
    out_name = par.get("output filename", required=True)
    S, sidx = _subset_keys(par)
    samples = _siddon_samples(par, "backprojector")
    if par.get("attenuation image filename") is not None:
        raise FormatError("'attenuation image filename' selects the attenuated "
                          "projectors, which are not on this path")
    ref_img, (dx, dy, dz) = read_image(out_image)
    nz, ny, nx = ref_img.shape
    ctx = hip.Context(_gpu_id(par))
    t0 = time.perf_counter()
    if itype == "Sinogram2D":
        rfov = atof(par.get("radius fov (in mm)", required=True))
        rscanner = atof(par.get("cylindrical pet radius (in mm)", required=True))
        s2 = read_sinogram2d(infile)
        if nz != 1 or nx != ny:
            raise FormatError("Sinogram2D input needs a square one-slice output image")
        plan = hip.Plan.for_2d(ctx, nx, dx, s2.shape[0], s2.shape[1], rfov, rscanner)
        sub = s2[sidx::S] if S > 0 else s2
        img = plan.backproject(torch.from_numpy(np.ascontiguousarray(sub[None])).cuda(),
                               subset=sidx, nsubsets=max(S, 1), mode=hip.BP_GATHER)
    elif itype in ("Sinogram3D", "Sinogram3DSiemensMmr"):
        sino = read_sinogram3d(infile)
        plan = _plan3d(ctx, itype, sino, par, ref_img.shape, (dx, dy, dz), samples)
        if itype == "Sinogram3DSiemensMmr" and S > 0:
            plan.set_scanner("mmr_subset")
        bins = sino["bins"][:, sidx::S, :] if S > 0 else sino["bins"]
        mode = hip.BP_SCATTER if samples != [1, 1] else hip.BP_GATHER
        img = plan.backproject(torch.from_numpy(np.ascontiguousarray(bins)).cuda(),
                               subset=sidx, nsubsets=max(S, 1), mode=mode)
    elif itype in ("Sinograms2D", "Sinograms2DinCylindrical3Dpet"):
        rfov, _, rscanner = _cyl(par)
        sinos, _ = read_image(infile)
        if sinos.shape[0] != nz:
            raise FormatError("backprojection of sinograms2D into an image with a different "
                              "number of slices")
        plan = multislice.slice_plan(ctx, nx, dx, sinos.shape[1], sinos.shape[2], rscanner,
                                     rfov=rfov)
        img = torch.from_numpy(np.stack([
            plan.backproject(torch.from_numpy(np.ascontiguousarray(sinos[k:k + 1])).cuda(),
                             mode=hip.BP_GATHER).cpu().numpy()[0] for k in range(nz)])).cuda()
    else:
        raise FormatError(f"input type '{itype}' is not available on this path")
    res = write_image(out_name, img.cpu().numpy().reshape(nz, ny, nx), dx, dy, dz)
    log(f"Backprojection finished. Processing time: {time.perf_counter() - t0}sec.")
    return res


def _block_size(par, key):
    """ProcessBlockSizeString (readCudaParameters.cpp:23-49): "{x,y,z}", the text
    between the last '{' and the last '}', split at blanks and commas.  Required, as
    there (:51-146)."""
    text = par.get(key, required=True)
    a, b = text.rfind("{"), text.rfind("}")
    if a < 0 or b < 0:
        raise FormatError(f"'{key}': no {{}} around the block size")
    parts = [t for t in re.split(r"[ ,]+", text[a + 1:b]) if t]
    if len(parts) < 3:
        raise FormatError(f"'{key}': the block size needs x, y and z")
    return tuple(atoi(t) for t in parts[:3])


def run_cumlem(par_path, log=print):
    """cmdGPU/src/cuMLEM.cpp:252-430: the GPU MLEM command's keys -- "projector block
    size", "backprojector block size", "pixel update block size" (required, "{x,y,z}")
    and "gpu id" -- and its dispatch: Sinogram3D input runs MLEM, Sinogram3DSiemensMmr
    input OSEM when "number of subsets" is given.  The reconstruction is the CPU
    reference's (MlemSinogram3d / OsemSinogram3d rules, the parity target); the block
    sizes are checked and logged, and the kernels keep their own launch shapes."""
    _check_par(par_path)
    par = ParFile(par_path, "MLEM")
    sizes = {k: _block_size(par, k) for k in ("projector block size", "backprojector block size",
                                               "pixel update block size")}
    for k, v in sizes.items():
        log(f"{k}: {v} (accepted; the gfx950 kernels use their own launch shapes)")
    itype = par.get("input type", required=True)
    if itype not in ("Sinogram3D", "Sinogram3DSiemensMmr"):
        raise FormatError(f"input type '{itype}' is not available to cuMLEM on this path")
    # cuMLEM.cpp:279-316: "CuSiddonProjector" names the GPU Siddon projector
    return run("MLEM", par_path, log=log, gpu=_gpu_id(par), mmr_osem_if_subsets=True,
               projectors=GPU_PROJECTORS)


COMMANDS = {"OSEM": lambda p: run("OSEM", p), "MLEM": lambda p: run("MLEM", p),
            "project": run_project, "backproject": run_backproject, "cuMLEM": run_cumlem}


def main(argv=None):
    argv = sys.argv[1:] if argv is None else argv
    if len(argv) != 2 or argv[0] not in COMMANDS:
        print("usage: python -m apirl_amd.cli OSEM|MLEM|project|backproject|cuMLEM "
              "parameters.par", file=sys.stderr)
        return 1
    try:
        COMMANDS[argv[0]](os.path.abspath(argv[1]))
    except (FormatError, hip.ApirlError, KeyError, ValueError, OSError) as e:
        print(f"{argv[0]}: {e}", file=sys.stderr)
        return 1
    return 0


if __name__ == "__main__":
    sys.exit(main())
