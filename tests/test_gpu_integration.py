"""Drop-in check: the reference's OWN EM drivers (Mlem2d, OsemSinogram3d, compiled
from its sources into oracle/_ref/) running with the MI355X projector plugged in
through its reference-side binding (integration/HipSiddonProjector.cpp), compared
with the same drivers running the reference's SiddonProjector (the goldens)."""
import json
import os
import sys
import tempfile

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRV = os.path.join(ROOT, "oracle", "_ref", "ref_hip_driver")

needs_driver = pytest.mark.skipif(not os.path.exists(DRV),
                                  reason="oracle/_ref/ref_hip_driver not built "
                                         "(needs the reference tree at build time)")


def run(cmd, **kw):
    """Runs the driver through the session's launcher: this process has already used
    the GPU, so it must not fork and exec the driver itself."""
    from conftest import LAUNCHER
    args = [DRV, cmd] + [f"{k}={v}" for k, v in kw.items()]
    print("case:", " ".join(args[1:3]), flush=True)
    rc, out, err = LAUNCHER.run(args, env=dict(os.environ, OMP_NUM_THREADS="1"), timeout=300)
    assert rc == 0, err[-2000:]
    assert "HipSiddonProjector" not in err, err[-2000:]
    return out


def bits(a):
    return np.ascontiguousarray(a, np.float32).ravel().view(np.int32)


GEOM3 = ["nx", "ny", "nz", "dx", "dz", "nproj", "nr", "rings", "rfov", "zfov", "rscanner", "span",
         "mrd"]


@needs_driver
def test_reference_project3d_with_hip_projector_bitexact(golden):
    f = golden("mid3d_project_backproject.npz")
    g = json.loads(str(f["geom"]))
    with tempfile.TemporaryDirectory() as td:
        f["img"].astype(np.float32).tofile(os.path.join(td, "img.f32"))
        run("proj3d", projector="hip", **{k: g[k] for k in GEOM3},
            img=os.path.join(td, "img.f32"), out=os.path.join(td, "p.f32"))
        got = np.fromfile(os.path.join(td, "p.f32"), np.float32)
    assert np.array_equal(bits(got), bits(f["proj"]))


@needs_driver
def test_reference_mlem2d_with_hip_gather_bitexact(golden):
    """Mlem2d (recon/src/Mlem2d.cpp) for 60 iterations with HipSiddonProjector as
    forward and backprojector: the final image equals the reference's bit for bit."""
    f = golden("c1_mlem2d.npz")
    g = json.loads(str(f["geom"]))
    with tempfile.TemporaryDirectory() as td:
        f["meas"].astype(np.float32).tofile(os.path.join(td, "meas.f32"))
        np.ones((128, 128), np.float32).tofile(os.path.join(td, "init.f32"))
        run("mlem2d", projector="hip", bp="gather", nx=g["nx"], dx=g["dx"], nproj=g["nproj"],
            nr=g["nr"], rscanner=g["rscanner"], iters=g["iters"],
            meas=os.path.join(td, "meas.f32"), init=os.path.join(td, "init.f32"), outdir=td)
        final = np.fromfile(os.path.join(td, "final.f32"), np.float32)
        sens = np.fromfile(os.path.join(td, "sens.f32"), np.float32)
        lik = np.fromfile(os.path.join(td, "lik.f32"), np.float32)
    assert np.array_equal(bits(sens), bits(f["sens"]))
    assert np.array_equal(bits(final), bits(f["final"]))
    assert np.array_equal(bits(lik), bits(f["lik"]))


@needs_driver
def test_reference_osem3d_with_hip_gather(golden):
    """OsemSinogram3d (recon/src/OsemSinogram3d.cpp), 7 subsets x 2 iterations, with
    HipSiddonProjector: projections are bit-exact, backprojection sums run in a fixed
    but different order, so the image matches within 1e-4 relative L2 (NaN where the
    reference has NaN)."""
    f = golden("mid3d_osem.npz")
    g = json.loads(str(f["geom"]))
    with tempfile.TemporaryDirectory() as td:
        f["meas"].astype(np.float32).tofile(os.path.join(td, "meas.f32"))
        np.ones((g["nz"], g["ny"], g["nx"]), np.float32).tofile(os.path.join(td, "init.f32"))
        run("osem3d", projector="hip", bp="gather", **{k: g[k] for k in GEOM3},
            subsets=g["subsets"], iters=g["iters"], meas=os.path.join(td, "meas.f32"),
            init=os.path.join(td, "init.f32"), outdir=td)
        final = np.fromfile(os.path.join(td, "final.f32"), np.float32)
    ref = f["final"].ravel()
    assert np.array_equal(np.isnan(final), np.isnan(ref))
    m = np.isfinite(ref)
    e = np.linalg.norm(final[m].astype(np.float64) - ref[m]) / np.linalg.norm(ref[m].astype(np.float64))
    print(f"osem3d via reference driver + HIP projector: rel L2 {e:.2e}", flush=True)
    assert e <= 1e-4


def mmr_header(td, golden):
    """The mMR sinogram of the mmr golden written as Interfile, with its geometry."""
    sys.path.insert(0, ROOT)
    from apirl_amd import hip, interfile
    f = golden("mmr.npz")
    g = json.loads(str(f["geom"]))
    nsino, mn, mx = hip.segment_table(g["rings"], g["span"], g["mrd"])
    meas = f["meas"].astype(np.float32).reshape(int(nsino.sum()), g["nproj"], g["nr"])
    hdr = interfile.write_sinogram3d(os.path.join(td, "mmr"), meas, nsino, mn, mx, g["rings"],
                                     8.18, 180.0 / g["nproj"])
    return f, g, hdr


def voxel_rel(a, b):
    a = np.asarray(a, np.float64).ravel()
    b = np.asarray(b, np.float64).ravel()
    return float(np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-6 * np.abs(b).max())))


@needs_driver
def test_reference_mmr_with_hip_projector(golden):
    """A Sinogram3DSiemensMmr read by the reference from Interfile, projected and
    backprojected through HipSiddonProjector (which takes the mMR LORs from the 2D
    sinograms' class): projection bit-exact, gather backprojection within 1e-5 per
    voxel, against the reference's SiddonProjector."""
    with tempfile.TemporaryDirectory() as td:
        f, g, hdr = mmr_header(td, golden)
        f["img"].astype(np.float32).tofile(os.path.join(td, "img.f32"))
        run("mmr3d", projector="hip", bp="gather", header=hdr, nx=g["nx"], ny=g["ny"],
            nz=g["nz"], dx=g["dx"], dz=g["dz"], img=os.path.join(td, "img.f32"),
            out_proj=os.path.join(td, "p.f32"), out_back=os.path.join(td, "b.f32"))
        proj = np.fromfile(os.path.join(td, "p.f32"), np.float32)
        back = np.fromfile(os.path.join(td, "b.f32"), np.float32)
    assert np.array_equal(bits(proj), bits(f["proj"]))
    assert voxel_rel(back, f["back"]) <= 1e-5


@needs_driver
def test_reference_mlem3d_on_mmr_with_hip_gather(golden):
    """MlemSinogram3d of the reference on the mMR sinogram, 3 iterations, with
    HipSiddonProjector (gather): sensitivity within 1e-5 per voxel and image within
    1e-4 relative L2 of the reference's run with its own projector."""
    m = golden("mmr_mlem3d.npz")
    with tempfile.TemporaryDirectory() as td:
        f, g, hdr = mmr_header(td, golden)
        np.ones((g["nz"], g["ny"], g["nx"]), np.float32).tofile(os.path.join(td, "init.f32"))
        run("mmr_mlem3d", projector="hip", bp="gather", header=hdr, nx=g["nx"], ny=g["ny"],
            nz=g["nz"], dx=g["dx"], dz=g["dz"], init=os.path.join(td, "init.f32"), iters=3,
            outdir=td)
        final = np.fromfile(os.path.join(td, "final.f32"), np.float32)
        sens = np.fromfile(os.path.join(td, "sens.f32"), np.float32)
    assert voxel_rel(sens, m["sens"]) <= 1e-5
    ref = m["final"].ravel().astype(np.float64)
    assert np.isfinite(final).all()
    e = np.linalg.norm(final - ref) / np.linalg.norm(ref)
    print(f"mMR MLEM3D via reference driver + HIP projector: rel L2 {e:.2e}", flush=True)
    assert e <= 1e-4


def run_refused(cmd, **kw):
    """Runs the driver expecting HipSiddonProjector to refuse the call: the reference's
    drivers ignore the projector's return value, so the refusal shows as its message."""
    from conftest import LAUNCHER
    args = [DRV, cmd] + [f"{k}={v}" for k, v in kw.items()]
    print("case:", " ".join(args[1:3]), flush=True)
    rc, out, err = LAUNCHER.run(args, env=dict(os.environ, OMP_NUM_THREADS="1"), timeout=300)
    return rc, err


@needs_driver
def test_reference_mmr_osem_with_hip_gather(golden):
    """OsemSinogram3d of the reference on the mMR sinogram with 3 subsets, with
    HipSiddonProjector: the subset sinograms the reference builds are cylindrical ones
    at the mMR's grids (Sinogram3DSiemensMmr::getSubset), which the binding takes
    from their class, so the image meets the bar against the reference's own run."""
    sys.path.insert(0, ROOT)
    from apirl_amd import hip, interfile
    c = golden("corrections.npz")
    g = json.loads(str(c["geom_mmr_osem"]))
    nsino, mn, mx = hip.segment_table(g["rings"], g["span"], g["mrd"])
    counts = np.random.default_rng(g["seed"]).poisson(
        2.0, (int(nsino.sum()), g["nproj"], g["nr"])).astype(np.float32)
    with tempfile.TemporaryDirectory() as td:
        hdr = interfile.write_sinogram3d(os.path.join(td, "mmro"), counts, nsino, mn, mx,
                                         g["rings"], 4.09, 180.0 / g["nproj"])
        np.ones((g["nz"], g["ny"], g["nx"]), np.float32).tofile(os.path.join(td, "init.f32"))
        run("mmr_osem3d", projector="hip", bp="gather", header=hdr, nx=g["nx"], ny=g["ny"],
            nz=g["nz"], dx=g["dx"], dz=g["dz"], init=os.path.join(td, "init.f32"),
            iters=g["iters"], subsets=g["subsets"], outdir=td)
        final = np.fromfile(os.path.join(td, "final.f32"), np.float32)
    ref = c["mmr_osem_final"].ravel()
    assert np.array_equal(np.isnan(final), np.isnan(ref))
    m = np.isfinite(ref)
    e = np.linalg.norm(final[m].astype(np.float64) - ref[m]) / np.linalg.norm(ref[m].astype(np.float64))
    print(f"mMR OSEM via reference driver + HIP projector: rel L2 {e:.2e}", flush=True)
    assert e <= 1e-4


@needs_driver
def test_reference_divide_backproject_with_hip_gather(golden):
    """Projector::DivideAndBackproject through the binding in gather mode: within
    1e-5 per voxel of the reference's, bit-identical from run to run."""
    f = golden("divide_backproject.npz")
    g = json.loads(str(f["geom3d"]))
    outs = []
    with tempfile.TemporaryDirectory() as td:
        f["meas3d"].astype(np.float32).tofile(os.path.join(td, "m.f32"))
        (f["est3d_x2"].astype(np.float32) * np.float32(0.5)).tofile(os.path.join(td, "e.f32"))
        for rep in range(2):
            out = os.path.join(td, f"d{rep}.f32")
            run("divback3d", projector="hip", bp="gather", **{k: g[k] for k in GEOM3},
                meas=os.path.join(td, "m.f32"), est=os.path.join(td, "e.f32"), out=out)
            outs.append(np.fromfile(out, np.float32))
    assert voxel_rel(outs[0], f["img3d"]) <= 1e-5
    assert np.array_equal(bits(outs[0]), bits(outs[1]))


@needs_driver
def test_reference_oversampled_with_hip_projector(golden):
    """HipSiddonProjector(gpu, mode, 3, 2), the counterpart of SiddonProjector(3, 2):
    projection bit-exact, backprojection within 1e-5 per voxel of the reference's with
    either backprojector (the gather takes a bin's six LORs in turn)."""
    f = golden("oversampled.npz")
    g = json.loads(str(f["geom"]))
    m = golden("mid3d_project_backproject.npz")
    geo = {k: g[k] for k in GEOM3}
    with tempfile.TemporaryDirectory() as td:
        m["img"].astype(np.float32).tofile(os.path.join(td, "img.f32"))
        m["meas"].astype(np.float32).tofile(os.path.join(td, "meas.f32"))
        run("proj3d", projector="hip", bp="scatter", ns=g["ns"], nas=g["nas"], **geo,
            img=os.path.join(td, "img.f32"), out=os.path.join(td, "p.f32"))
        run("back3d", projector="hip", bp="scatter", ns=g["ns"], nas=g["nas"], **geo,
            sino=os.path.join(td, "meas.f32"), out=os.path.join(td, "b.f32"))
        proj = np.fromfile(os.path.join(td, "p.f32"), np.float32)
        back = np.fromfile(os.path.join(td, "b.f32"), np.float32)
        run("back3d", projector="hip", bp="gather", ns=g["ns"], nas=g["nas"], **geo,
            sino=os.path.join(td, "meas.f32"), out=os.path.join(td, "b2.f32"))
        back_g = np.fromfile(os.path.join(td, "b2.f32"), np.float32)
    assert np.array_equal(bits(proj), bits(f["proj"]))
    assert voxel_rel(back, f["back"]) <= 1e-5
    assert voxel_rel(back_g, f["back"]) <= 1e-5


@needs_driver
def test_reference_multiple_lors_per_bin_refused(golden):
    """setMultipleLorsPerBin(true) (Projector.h:60): the binding traces each bin on its
    sinogram's mean ring pair only, so it returns false and says why instead of
    returning mean-pair results."""
    f = golden("mid3d_project_backproject.npz")
    g = json.loads(str(f["geom"]))
    with tempfile.TemporaryDirectory() as td:
        f["img"].astype(np.float32).tofile(os.path.join(td, "img.f32"))
        rc, err = run_refused("proj3d", projector="hip", multilor=1, **{k: g[k] for k in GEOM3},
                              img=os.path.join(td, "img.f32"), out=os.path.join(td, "p.f32"))
    assert "useMultipleLorsPerBin" in err
