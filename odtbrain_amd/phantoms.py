"""Seeded analytic phantoms and their Rytov sinograms.

Synthetic inputs for tests and benchmarks: refractive-index phantoms made of
spheres/ellipsoids whose projections are computed analytically (chord
lengths), so no wave solver is needed.  Geometry follows the reference's
convention for rotation about the y axis (odtbrain/_alg3d_bpp.py:348-350):
at angle phi the detector axis is t = (cos phi, 0, sin phi) and the light
travels along s0 = (-sin phi, 0, cos phi); rotation centre at the array
centre ((N-1)/2 in each axis).
"""
import numpy as np


def _chord_ellipsoid(p0, s, center, q):
    """Chord length of rays p0 + t*s (|s| = 1) through x^T q x <= 1."""
    d = p0 - center
    qs = s @ q
    a = qs @ s
    b = d @ qs
    c = np.einsum("...i,ij,...j->...", d, q, d) - 1.0
    disc = b * b - a * c
    return np.where(disc > 0, 2.0 * np.sqrt(np.maximum(disc, 0)) / a, 0.0)


def _rays_y_axis(phi, ny, nx):
    """Ray origins (detector plane through the centre) and direction."""
    t = np.array([np.cos(phi), 0.0, np.sin(phi)])
    s = np.array([-np.sin(phi), 0.0, np.cos(phi)])
    xd = np.arange(nx) - (nx - 1) / 2.0
    yd = np.arange(ny) - (ny - 1) / 2.0
    p0 = xd[None, :, None] * t[None, None, :] + \
        yd[:, None, None] * np.array([0.0, 1.0, 0.0])[None, None, :]
    return p0, s


def sphere_sinogram_3d(N=64, A=100, radius=16.0, n_sphere=1.343, nm=1.333,
                       res=6.0, noise=0.01, seed=0):
    """Rytov sinogram of a centred sphere (complex128, shape (A, N, N)).

    u = exp(i k0 (n - nm) chord) + complex Gaussian noise (sigma on real and
    imaginary part), Rytov data uB = log(u); k0 = 2 pi / res.
    """
    rng = np.random.default_rng(seed)
    angles = np.linspace(0, 2 * np.pi, A, endpoint=False)
    c = (N - 1) / 2.0
    y, x = np.mgrid[0:N, 0:N]
    rho2 = (x - c) ** 2 + (y - c) ** 2
    chord = 2.0 * np.sqrt(np.maximum(radius ** 2 - rho2, 0.0))
    k0 = 2 * np.pi / res
    u = np.exp(1j * k0 * (n_sphere - nm) * chord)[None].repeat(A, axis=0)
    u = u + noise * (rng.standard_normal(u.shape) +
                     1j * rng.standard_normal(u.shape))
    return np.log(u), angles


def random_ellipsoids(count, semi=(20, 60), dn=(0.005, 0.02), spread=30.0,
                      seed=0):
    """Seeded ellipsoids: (center, quadratic form, delta n) triples."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        axes = rng.uniform(*semi, size=3)
        qmat, _r = np.linalg.qr(rng.standard_normal((3, 3)))
        q = qmat @ np.diag(1.0 / axes ** 2) @ qmat.T
        center = rng.uniform(-spread, spread, size=3)
        out.append((center, q, rng.uniform(*dn)))
    return out


def ellipsoid_sinogram_3d(N=256, A=360, count=3, res=6.0, phase_noise=0.02,
                          seed=1, dtype=np.complex64, semi=(20, 60),
                          spread=30.0):
    """Rytov sinogram uB = i*phase of seeded overlapping ellipsoids.

    phase = k0 * sum_j dn_j * chord_j + N(0, phase_noise) (radians);
    returns (uB (A, N, N), angles (A,)).
    """
    rng = np.random.default_rng(seed)
    ells = random_ellipsoids(count, semi=semi, spread=spread, seed=seed)
    angles = np.linspace(0, 2 * np.pi, A, endpoint=False)
    k0 = 2 * np.pi / res
    out = np.empty((A, N, N), dtype=dtype)
    for i, phi in enumerate(angles):
        p0, s = _rays_y_axis(phi, N, N)
        ph = np.zeros((N, N))
        for center, q, dn in ells:
            ph += dn * _chord_ellipsoid(p0, s, center, q)
        ph = k0 * ph + rng.normal(0.0, phase_noise, size=(N, N))
        out[i] = 1j * ph
    return out, angles


def tilted_sphere_sinogram(N=192, A=240, count=5, coverage=1.4 * np.pi,
                           tilt=0.2, res=6.0, seed=2, dtype=np.complex64):
    """Rytov sinogram of seeded spheres for a tilted rotation axis.

    Angles: ``linspace(0, coverage, A, endpoint=False)`` plus seeded jitter
    of +-0.4 steps, sorted.  The propagation direction of projection i is
    the great-circle point of the tilted axis [0, cos(tilt), sin(tilt)]
    (first point [0, 0, 1]); detector axes are orthonormal to it.
    Returns (uB (A, N, N), angles (A,), tilted_axis).
    """
    from ._geometry import sphere_points_from_angles_and_tilt
    rng = np.random.default_rng(seed)
    step = coverage / A
    angles = np.sort(np.linspace(0, coverage, A, endpoint=False) +
                     rng.uniform(-0.4, 0.4, A) * step)
    axis = np.array([0.0, np.cos(tilt), np.sin(tilt)])
    pts = sphere_points_from_angles_and_tilt(angles, axis)
    spheres = [(rng.uniform(-40, 40, size=3), rng.uniform(10, 40),
                rng.uniform(0.003, 0.015)) for _ in range(count)]
    k0 = 2 * np.pi / res
    xd = np.arange(N) - (N - 1) / 2.0
    out = np.empty((A, N, N), dtype=dtype)
    for i in range(A):
        s = pts[i] / np.linalg.norm(pts[i])
        e1 = np.cross(axis, s)
        e1 /= np.linalg.norm(e1)
        e2 = np.cross(s, e1)
        p0 = xd[None, :, None] * e1 + xd[:, None, None] * e2
        ph = np.zeros((N, N))
        for center, r, dn in spheres:
            d = p0 - center
            along = d @ s
            rho2 = np.einsum("...i,...i->...", d, d) - along ** 2
            ph += dn * 2.0 * np.sqrt(np.maximum(r * r - rho2, 0.0))
        out[i] = 1j * k0 * ph
    return out, angles, axis


def disk_sinogram_2d(N=256, A=180, radius=None, dn=0.01, nm=1.333, res=6.0,
                     offset=(0.0, 0.0), seed=0):
    """Rytov sinogram (complex64, shape (A, N)) of a disk, uB = i*phase.

    The disk (index nm + dn, centre offset from the rotation centre) is
    projected analytically at angles uniform on [0, 2 pi); seeded phase
    noise of 0.01 rad."""
    rng = np.random.default_rng(seed)
    r = N / 6.0 if radius is None else radius
    phi = np.linspace(0, 2 * np.pi, A, endpoint=False)
    xd = np.arange(N) - (N - 1) / 2.0
    k0 = 2 * np.pi / res
    # detector coordinate of the disk centre at each angle
    cx, cz = offset
    xc = cx * np.cos(phi) + cz * np.sin(phi)
    rho2 = (xd[None, :] - xc[:, None]) ** 2
    chord = 2 * np.sqrt(np.maximum(r * r - rho2, 0.0))
    phase = k0 * dn * chord + rng.normal(0, 0.01, chord.shape)
    return (1j * phase).astype(np.complex64), phi


def batch_jobs(scale=1.0, seed=0):
    """The batched workload: 32 2D sinograms (N in {128, 256, 384, 512,
    1024}, 90-720 angles) and 16 3D sinograms (N 64-160, 60-300 angles),
    seeded disk/sphere phantoms with dn in [0.005, 0.02], complex64.
    Returns a list of (kind, uSin, angles, res, nm).  ``scale`` < 1 shrinks
    the angle counts (for tests)."""
    rng = np.random.default_rng(seed)
    jobs = []
    sizes2 = [128, 256, 384, 512, 1024]
    for i in range(32):
        N = sizes2[i % len(sizes2)]
        A = max(8, int(rng.integers(90, 721) * scale))
        dn = float(rng.uniform(0.005, 0.02))
        s, ang = disk_sinogram_2d(N=N, A=A, dn=dn, seed=seed + i)
        jobs.append(("2d", s, ang, 6.0, 1.333))
    for i in range(16):
        N = int(rng.integers(64, 161))
        A = max(8, int(rng.integers(60, 301) * scale))
        r = float(rng.uniform(0.15, 0.3) * N)
        n_s = 1.333 + float(rng.uniform(0.005, 0.02))
        s, ang = sphere_sinogram_3d(N=N, A=A, radius=r, n_sphere=n_s,
                                    seed=seed + 100 + i)
        jobs.append(("3d", s.astype(np.complex64), ang, 6.0, 1.333))
    return jobs
