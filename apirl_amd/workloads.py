"""Seeded reconstruction jobs of uneven size for the batched and served workloads.

Job k (seed k) draws uniformly one of four geometries:
  2d_128  : 2D, 128x128 image @ 2.0 mm, 192 radial bins x 192 angles, scanner radius
            400 mm (the 2D MLEM geometry);
  2d_256  : 2D, 256x256 image @ 2.0445 mm, 344 radial bins x 252 angles;
  3d_small: 3D, 128x128x47 @ 2.0 mm, 192 x 192 x 233 sinograms (24 rings, span 5,
            maximum ring difference 14; the nearest reachable count to 235);
  3d_c2   : 3D, 256x256x81 @ 2.0445 / 2.027 mm, 344 x 252 x 837 sinograms (64 rings,
            span 11, maximum ring difference 60);
a phantom of 1-8 random ellipsoids (values in [0.5, 4]), Poisson counts drawn
log-uniformly in [1e5, 1e8], 1-21 subsets and 1-10 iterations, and optionally a
multiplicative (attenuation) sinogram with values in [1, 20] that enters the
sensitivity only, as in the reference (MlemSinogram3d.cpp:312-331)."""
import numpy as np

from . import hip, phantoms

GEOMETRIES = {
    "2d_128": dict(dims=2, nx=128, dx=2.0, nproj=192, nr=192, rfov=128.0, rscanner=400.0),
    "2d_256": dict(dims=2, nx=256, dx=2.0445, nproj=252, nr=344, rfov=261.696, rscanner=400.0),
    "3d_small": dict(dims=3, nx=128, ny=128, nz=47, dx=2.0, dz=2.0, nproj=192, nr=192,
                     rings=24, rfov=128.0, zfov=94.0, rscanner=400.0, span=5, mrd=14),
    "3d_c2": dict(dims=3, nx=256, ny=256, nz=81, dx=2.0445, dz=2.027, nproj=252, nr=344,
                  rings=64, rfov=261.696, zfov=164.187, rscanner=400.0, span=11, mrd=60),
}
SMALL = ("2d_128", "2d_256", "3d_small")


def make_plan(ctx, name):
    g = GEOMETRIES[name]
    if g["dims"] == 2:
        return hip.Plan.for_2d(ctx, g["nx"], g["dx"], g["nproj"], g["nr"], g["rfov"],
                               g["rscanner"])
    return hip.Plan.for_3d(ctx, g["nx"], g["ny"], g["nz"], g["dx"], g["dz"], g["nproj"],
                           g["nr"], g["rfov"], g["rscanner"], g["rings"], g["zfov"], g["span"],
                           g["mrd"])


def draw(seed, kinds=tuple(GEOMETRIES), attenuation=False, max_iterations=10):
    """The parameters of job `seed` (no device work)."""
    rng = np.random.default_rng(seed)
    kind = kinds[int(rng.integers(len(kinds)))]
    g = GEOMETRIES[kind]
    return dict(seed=seed, kind=kind, subsets=int(rng.integers(1, min(21, g["nproj"]) + 1)),
                iterations=int(rng.integers(1, max_iterations + 1)),
                counts=float(10 ** rng.uniform(5, 8)),
                attenuation=bool(attenuation and rng.integers(2)))


def materialise(spec, plan, torch):
    """Device tensors of a job: measured sinogram (Poisson), initial image, output
    image and, when drawn, the multiplicative sinogram."""
    g = GEOMETRIES[spec["kind"]]
    rng = np.random.default_rng(1000003 + spec["seed"])
    nz = 1 if g["dims"] == 2 else g["nz"]
    img = phantoms.ellipsoids3d(g["nx"], nz, g["dx"], g.get("dz", 1.0), rng)
    x = torch.from_numpy(img.reshape(plan.image_shape())).cuda()
    truth = plan.project(x)
    gen = torch.Generator(device="cuda")
    gen.manual_seed(spec["seed"])
    s = float(truth.double().sum())
    meas = torch.poisson(truth * (spec["counts"] / s if s > 0 else 0.0), generator=gen)
    mult = None
    if spec["attenuation"]:
        mult = torch.empty_like(meas).uniform_(1.0, 20.0, generator=gen)
    init = torch.ones(plan.image_shape(), dtype=torch.float32, device="cuda")
    out = torch.empty_like(init)
    return dict(meas=meas.contiguous(), mult=mult, init=init, out=out)


def work_units(spec):
    """Ray-voxel work of a job in units of its sinogram bins x iterations (for
    reporting the mix; the measured figure is jobs/s)."""
    g = GEOMETRIES[spec["kind"]]
    nsino = 1 if g["dims"] == 2 else hip.ring_table(g["rings"], g["zfov"], g["span"],
                                                     g["mrd"])[0].size
    return nsino * g["nproj"] * g["nr"] * spec["iterations"]
