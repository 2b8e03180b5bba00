"""Synthetic activity images and count data for the workloads this library is
measured on.  Voxel centres follow the reference's grids: x_i = -R + dx (i + 1/2)
(data/src/Images.cpp:435-441, rows grow with +y) and, axially, the Siddon grid
whose slice k spans [dz/2 + k dz, dz/2 + (k+1) dz) (recon/src/Siddon.cpp:162,179),
i.e. centre z_k = (k + 1) dz."""
import numpy as np


def voxel_centres(nx, ny, nz, dx, dz):
    R = np.float32(dx) * nx / 2
    x = -R + np.float32(dx) * (np.arange(nx, dtype=np.float32) + 0.5)
    y = -R + np.float32(dx) * (np.arange(ny, dtype=np.float32) + 0.5)
    z = np.float32(dz) * (np.arange(nz, dtype=np.float32) + 1.0)
    return x, y, z


def disc2d(nx, dx, radius=100.0, value=1.0, hot_centre=(30.0, 20.0), hot_diameter=10.0,
           hot_value=4.0):
    """Uniform disc plus a hot disc (2D MLEM workload, 128x128 @ 2 mm)."""
    x, y, _ = voxel_centres(nx, nx, 1, dx, 1.0)
    X, Y = np.meshgrid(x, y)
    img = np.where(X * X + Y * Y <= radius * radius, value, 0.0).astype(np.float32)
    hx, hy = hot_centre
    hr = hot_diameter / 2
    img[(X - hx) ** 2 + (Y - hy) ** 2 <= hr * hr] = hot_value
    return img


def cylinder3d(nx, nz, dx, dz, radius=100.0, value=1.0):
    x, y, _ = voxel_centres(nx, nx, nz, dx, dz)
    X, Y = np.meshgrid(x, y)
    sl = np.where(X * X + Y * Y <= radius * radius, value, 0.0).astype(np.float32)
    return np.repeat(sl[None], nz, axis=0)


def cylinder_points3d(nx, nz, dx, dz, radius=100.0, npoints=6, point_value=50.0, seed=0):
    """Cylinder (value 1) filling the axial FOV plus single-voxel point sources at
    seeded random voxels inside it (3D OSEM workload, 256x256x81)."""
    img = cylinder3d(nx, nz, dx, dz, radius)
    rng = np.random.default_rng(seed)
    inside = np.argwhere(img[0] > 0)
    pts = []
    for _ in range(npoints):
        j, i = inside[rng.integers(len(inside))]
        k = int(rng.integers(nz))
        img[k, j, i] = point_value
        pts.append((k, j, i))
    return img, pts


def cylinder_spheres3d(nx, nz, dx, dz, radius=100.0, diameters=(10, 13, 17, 22, 28, 37),
                       contrast=4.0, ring_radius=57.2):
    """Cylinder with hot spheres on a ring (precision-study workload)."""
    img = cylinder3d(nx, nz, dx, dz, radius)
    x, y, z = voxel_centres(nx, nx, nz, dx, dz)
    Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
    zc = z[nz // 2]
    for n, d in enumerate(diameters):
        a = 2 * np.pi * n / len(diameters)
        cx, cy = ring_radius * np.cos(a), ring_radius * np.sin(a)
        m = (X - cx) ** 2 + (Y - cy) ** 2 + (Z - zc) ** 2 <= (d / 2) ** 2
        img[m] = contrast
    return img


def ellipsoids3d(nx, nz, dx, dz, rng, nmin=1, nmax=8, vmin=0.5, vmax=4.0):
    """Random ellipsoid mixture inside the transaxial FOV (batched-jobs workload)."""
    x, y, z = voxel_centres(nx, nx, nz, dx, dz)
    Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
    R = dx * nx / 2
    img = np.zeros((nz, nx, nx), np.float32)
    for _ in range(int(rng.integers(nmin, nmax + 1))):
        c = rng.uniform(-0.5 * R, 0.5 * R, 2)
        cz = rng.uniform(z[0], z[-1]) if nz > 1 else z[0]
        ax = rng.uniform(0.1 * R, 0.45 * R, 2)
        az = rng.uniform(0.2, 0.6) * max(z[-1] - z[0], dz) if nz > 1 else 1e9
        m = ((X - c[0]) / ax[0]) ** 2 + ((Y - c[1]) / ax[1]) ** 2 + ((Z - cz) / az) ** 2 <= 1
        img[m] += np.float32(rng.uniform(vmin, vmax))
    return img


def poisson_counts(truth, total, rng):
    """Scale a noiseless projection to `total` expected counts and Poisson-sample it."""
    t = np.asarray(truth, np.float64)
    s = t.sum()
    lam = t * (total / s) if s > 0 else t
    return rng.poisson(lam).astype(np.float32)


def correction_sinograms(nsino, nang, nr, negative=False):
    """Deterministic additive and multiplicative sinograms [nsino][nang][nr] whose
    values vary with the angle, so a mix-up of angles between a subset and the full
    sinogram changes the result.  additive: 0.05..0.4 counts; with negative=True about
    one bin in 211 is -50, far below any projection there, so est + additive < 0 (the
    reference adds it without a clamp, Sinogram3D.cpp:857-873).  multiplicative (the
    attenuation-like factor of the sensitivity): 1..3."""
    q, a, r = np.meshgrid(np.arange(nsino, dtype=np.int64), np.arange(nang, dtype=np.int64),
                          np.arange(nr, dtype=np.int64), indexing="ij")
    add = 0.05 + 0.06 * (a % 5) + 0.002 * (r % 17) + 0.001 * (q % 3)
    if negative:
        add = np.where((q * 131 + a * 17 + r * 7) % 211 == 0, -50.0, add)
    mult = 1.0 + ((q * 7 + a * 3 + r) % 13) / 6.0
    return add.astype(np.float32), mult.astype(np.float32)
