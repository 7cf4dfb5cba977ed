"""Seeded synthetic initial conditions for the stage-update workloads.

Every field is a deterministic function of (seed, resolution): the random
Fourier coefficients are drawn once, in a fixed order over the mode cube
[-KMAX, KMAX]^3, with numpy's PCG64 generator, and the field is evaluated by
a separable direct sum (no FFT), so the same seed gives the same continuous
field at every resolution and on every host.

Layouts follow the reference's data model (Src/Array.hpp:476-479, i fastest):
``vc[nv, k, j, i]`` cell-centred primitives in the reference's variable
order (MHD: rho, vx1, vx2, vx3, Bx1, Bx2, Bx3, prs -- Src/MHD/mhd.hpp:21-41;
HD: rho, vx1, vx2, vx3, prs -- Src/HD/hd.hpp:21-34) and ``vs[d, k, j, i]``
the face-normal field on the lower face of cell (i,j,k) in direction d
(staggered index i means face i-1/2, reference CHANGES:18).  Boxes are
periodic, so one face per cell is stored.
"""
from __future__ import annotations

import numpy as np

KMAX = 8


def _modes(kmax: int = KMAX):
    r = np.arange(-kmax, kmax + 1)
    kx, ky, kz = np.meshgrid(r, r, r, indexing="ij")
    k2 = kx * kx + ky * ky + kz * kz
    return kx, ky, kz, k2


def _coefficients(rng: np.random.Generator, slope: float, kmax: int = KMAX):
    """Complex mode amplitudes with E|c_k|^2 = |k|^slope on 1 <= |k| <= kmax."""
    kx, ky, kz, k2 = _modes(kmax)
    g = rng.standard_normal((2,) + k2.shape)
    mask = (k2 >= 1) & (k2 <= kmax * kmax)
    amp = np.zeros(k2.shape)
    amp[mask] = np.power(k2[mask].astype(np.float64), 0.5 * slope)
    c = (g[0] + 1j * g[1]) * amp / np.sqrt(2.0)
    # variance of Re(sum_k c_k e^{i k.x}) is sum_k E|c_k|^2 / 2
    var = 0.5 * np.sum(amp * amp)
    return c, var


def _synth(c: np.ndarray, n, shift, kmax: int = KMAX) -> np.ndarray:
    """Evaluate Re sum_k c_k exp(2 pi i k.x) at x_a = (a + shift)/n per axis.

    Returns an array indexed [k, j, i] (x fastest).  ``n`` and ``shift`` are
    3-tuples (x, y, z); an axis with n == 1 is evaluated at x = 0 only for
    the zero mode (the box is flat in that direction).
    """
    r = np.arange(-kmax, kmax + 1)
    ex = [np.exp(2j * np.pi * np.outer((np.arange(n[d]) + shift[d]) / n[d], r))
          for d in range(3)]
    # c[kx, ky, kz]
    t1 = np.einsum("xyz,cz->xyc", c, ex[2])           # [kx, ky, z]
    t2 = np.einsum("xyc,by->cbx", t1, ex[1])          # [z, y, kx]
    t2 = t2.reshape(-1, t2.shape[2])
    # Re(t2 . ex0^T) as two real matrix products, written in [z*y, x] order
    out = np.ascontiguousarray(t2.real) @ np.ascontiguousarray(ex[0].real.T)
    out -= np.ascontiguousarray(t2.imag) @ np.ascontiguousarray(ex[0].imag.T)
    return out.reshape(n[2], n[1], n[0])


def _flat_modes(c: np.ndarray, dims: int) -> np.ndarray:
    """Keep only modes with zero wavenumber along the absent directions."""
    if dims == 3:
        return c
    c = c.copy()
    k = np.arange(-KMAX, KMAX + 1)
    if dims <= 2:
        c[:, :, k != 0] = 0.0
    if dims <= 1:
        c[:, k != 0, :] = 0.0
    return c


def _curl(ax, ay, az, n, dims):
    """Face field B = curl A of an edge-centred potential on [0,1]^dims with
    periodic forward differences (the discrete curl of Src/vector_curl.cpp:109);
    arrays are [k, j, i]."""
    nx, ny, nz = n
    idx, idy, idz = float(nx), float(ny), float(nz)  # 1/dx on [0,1]
    dj = lambda a: np.roll(a, -1, axis=1) - a
    di = lambda a: np.roll(a, -1, axis=2) - a
    dk = lambda a: np.roll(a, -1, axis=0) - a
    bx = dj(az) * idy
    by = -di(az) * idx
    bz = di(ay) * idx - dj(ax) * idy
    if dims == 3:
        bx = bx - dk(ay) * idz
        by = by + dk(ax) * idz
    return bx, by, bz


def mhd_turbulence(n, seed: int, gamma: float = 5.0 / 3.0, beta: float = 10.0,
                   mach: float = 0.1, drho: float = 0.1):
    """Random-field MHD box: the 3D/2D ideal-MHD workload.

    rho = 1 + drho*G (G unit-variance Gaussian random field, per-mode power
    |k|^-2, 1 <= |k| <= 8), p = 1, velocity components independent random
    fields scaled to rms |v| = mach * c_s (c_s = sqrt(gamma p / rho0)), face B
    the discrete curl of an edge-centred random vector potential with the same
    spectrum, scaled so that 2 <p> / <B^2> = beta.

    ``n`` is (nx, ny, nz); nz == 1 gives a 2D box (nz == ny == 1 is not an MHD
    workload here).  Returns (vc[8, nz, ny, nx], vs[dims, nz, ny, nx]).
    """
    nx, ny, nz = n
    dims = 3 if nz > 1 else 2
    rng = np.random.Generator(np.random.PCG64(seed))
    fields = []
    for _ in range(7):  # rho, vx, vy, vz, Ax, Ay, Az (fixed draw order)
        c, var = _coefficients(rng, -2.0)
        fields.append((_flat_modes(c, dims), var))
    cen = (0.5, 0.5, 0.5 if dims == 3 else 0.0)
    rho_c, rho_var = fields[0]
    # flat-mode filtering changes the variance; use the empirical one
    g = _synth(rho_c, n, cen)
    rho = 1.0 + drho * g / max(np.std(g), 1e-300)
    cs = np.sqrt(gamma * 1.0 / 1.0)
    vel = []
    for d in range(3):
        vc_, _ = fields[1 + d]
        vel.append(_synth(vc_, n, cen))
    vrms = np.sqrt(sum(np.mean(v * v) for v in vel))
    vel = [v * (mach * cs / vrms) for v in vel]
    # edge-centred potential: A_x at (x_i, y_{j-1/2}, z_{k-1/2}), etc.
    zf = 0.0
    zc = 0.5 if dims == 3 else 0.0
    ax = _synth(fields[4][0], n, (0.5, 0.0, zf))
    ay = _synth(fields[5][0], n, (0.0, 0.5, zf))
    az = _synth(fields[6][0], n, (0.0, 0.0, zc))

    curl = lambda ax, ay, az: _curl(ax, ay, az, n, dims)
    bx, by, bz = curl(ax, ay, az)
    bcx = 0.5 * (bx + np.roll(bx, -1, axis=2))
    bcy = 0.5 * (by + np.roll(by, -1, axis=1))
    bcz = 0.5 * (bz + np.roll(bz, -1, axis=0)) if dims == 3 else bz
    b2 = np.mean(bcx * bcx + bcy * bcy + bcz * bcz)
    s = np.sqrt(2.0 * 1.0 / beta / b2)
    bx, by, bz = curl(ax * s, ay * s, az * s)
    # in 2D only Bx, By are staggered; Bz stays cell-centred (Src/MHD/CT/ct.hpp:83-89)
    vs = np.stack([bx, by, bz]) if dims == 3 else np.stack([bx, by])
    bcx = 0.5 * (bx + np.roll(bx, -1, axis=2))
    bcy = 0.5 * (by + np.roll(by, -1, axis=1))
    bcz = 0.5 * (bz + np.roll(bz, -1, axis=0)) if dims == 3 else bz
    prs = np.ones_like(rho)
    vc = np.stack([rho, vel[0], vel[1], vel[2], bcx, bcy, bcz, prs])
    return np.ascontiguousarray(vc), np.ascontiguousarray(vs)


def rmhd_turbulence(n, seed: int, gamma: float = 4.0 / 3.0, lorentz_max: float = 3.0,
                    sigma: float = 1.0, drho: float = 0.5):
    """Random-field relativistic MHD box (the 3D RMHD workload's generator).

    rho = exp(drho*G) (log-normal), p/rho = 10**clip(-0.5 + 0.5*G', -2, 1)
    (so p/rho spans [0.01, 10]), spatial four-velocity u = Gamma v from random
    modes scaled so that the largest Lorentz factor is ``lorentz_max``, and a
    face field curl A scaled so that <B^2>/<rho h> = sigma.  G, G' are
    unit-variance Gaussian random fields with |k|^-2 power on 1 <= |k| <= 8.
    Returns (vc[8, nz, ny, nx], vs[3, nz, ny, nx])."""
    nx, ny, nz = n
    dims = 3 if nz > 1 else 2
    rng = np.random.Generator(np.random.PCG64(seed))
    fields = [_flat_modes(_coefficients(rng, -2.0)[0], dims) for _ in range(8)]
    cen = (0.5, 0.5, 0.5 if dims == 3 else 0.0)
    unit = lambda f: (lambda g: g / max(np.std(g), 1e-300))(_synth(f, n, cen))
    rho = np.exp(drho * unit(fields[0]))
    theta = 10.0 ** np.clip(-0.5 + 0.5 * unit(fields[1]), -2.0, 1.0)
    prs = rho * theta
    u4 = [unit(fields[2 + d]) for d in range(3)]
    umag = np.sqrt(sum(u * u for u in u4))
    s = np.sqrt(lorentz_max ** 2 - 1.0) / np.max(umag)
    u4 = [u * s for u in u4]
    lor = np.sqrt(1.0 + sum(u * u for u in u4))
    vel = [u / lor for u in u4]
    zc = 0.5 if dims == 3 else 0.0
    ax = _synth(fields[5], n, (0.5, 0.0, 0.0))
    ay = _synth(fields[6], n, (0.0, 0.5, 0.0))
    az = _synth(fields[7], n, (0.0, 0.0, zc))
    bx, by, bz = _curl(ax, ay, az, n, dims)
    cell = lambda bx, by, bz: (0.5 * (bx + np.roll(bx, -1, axis=2)),
                               0.5 * (by + np.roll(by, -1, axis=1)),
                               0.5 * (bz + np.roll(bz, -1, axis=0)) if dims == 3 else bz)
    bcx, bcy, bcz = cell(bx, by, bz)
    h = 1.0 + gamma / (gamma - 1.0) * theta
    scale = np.sqrt(sigma * np.mean(rho * h) / np.mean(bcx * bcx + bcy * bcy + bcz * bcz))
    bx, by, bz = _curl(ax * scale, ay * scale, az * scale, n, dims)
    bcx, bcy, bcz = cell(bx, by, bz)
    vs = np.stack([bx, by, bz]) if dims == 3 else np.stack([bx, by])
    vc = np.stack([rho, vel[0], vel[1], vel[2], bcx, bcy, bcz, prs])
    return np.ascontiguousarray(vc), np.ascontiguousarray(vs)


def hd_riemann(n: int, seed: int):
    """Seeded 1D shock tube on [0,1]: jump at x = 0.5, rho and p on each side
    drawn from U[0.1, 1), v = 0.  Returns vc[5, 1, 1, n]."""
    rng = np.random.Generator(np.random.PCG64(seed))
    rl, rr, pl, pr = rng.uniform(0.1, 1.0, size=4)
    x = (np.arange(n) + 0.5) / n
    left = x < 0.5
    vc = np.zeros((5, 1, 1, n))
    vc[0, 0, 0] = np.where(left, rl, rr)
    vc[4, 0, 0] = np.where(left, pl, pr)
    return vc


def mhd_batch_boxes(seed: int, nbox: int = 64, nmin: int = 32, nmax: int = 512):
    """Sizes and plasma betas of the batched 2D workload: N_b = floor(nmin *
    (nmax/nmin)**u_b), u_b ~ U[0,1) (log-uniform in [nmin, nmax)), beta_b ~
    U[1, 100].  Box b's IC is mhd_turbulence((N_b, N_b, 1), box_seed(seed, b),
    beta=beta_b)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    u = rng.uniform(0.0, 1.0, size=nbox)
    n = np.floor(nmin * (nmax / nmin) ** u).astype(int)
    beta = rng.uniform(1.0, 100.0, size=nbox)
    return [int(x) for x in n], [float(x) for x in beta]


def box_seed(seed: int, b: int) -> int:
    return 1000003 * (seed + 1) + b


def hd_blast(n, seed: int, gamma: float = 1.4):
    """Seeded HD box (1D, 2D or 3D, n = (nx, ny, nz)) on the unit cube: cell
    noise of 25% on rho and p around 1, random velocities of rms ~0.2, and a
    pressure of 10 inside radius 0.2 of a random centre, so the run has strong
    shocks, contacts and limiter switches everywhere.  Returns vc[5, nz, ny, nx]."""
    nx, ny, nz = n
    rng = np.random.Generator(np.random.PCG64(seed))
    shape = (nz, ny, nx)
    vc = np.empty((5,) + shape)
    vc[0] = 1.0 + 0.25 * rng.uniform(-1.0, 1.0, size=shape)
    for d in range(3):
        vc[1 + d] = 0.35 * rng.uniform(-1.0, 1.0, size=shape) if (nx, ny, nz)[d] > 1 else 0.0
    vc[4] = 1.0 + 0.25 * rng.uniform(-1.0, 1.0, size=shape)
    c = rng.uniform(0.3, 0.7, size=3)
    z, y, x = np.meshgrid((np.arange(nz) + 0.5) / nz, (np.arange(ny) + 0.5) / ny,
                          (np.arange(nx) + 0.5) / nx, indexing="ij")
    r2 = (x - c[0]) ** 2 + (y - c[1]) ** 2 * (ny > 1) + (z - c[2]) ** 2 * (nz > 1)
    vc[4][r2 < 0.04] = 10.0
    return vc


def hd_vacuum(n, vmag: float, prs: float):
    """Two streams leaving x = 0.5 at speed vmag (1D or 2D, n = (nx, ny, 1)):
    rho = 1 + 0.2 cos(2 pi y), p = prs, v1 = -vmag / +vmag left / right of
    the centre, and in 2D a transverse v2 = vmag/2 sin(2 pi y) near x = 0.5.
    The centre empties, so ConsToPrim meets negative pressures and energies
    (its small_prs floor).  Returns vc[5, 1, ny, nx]."""
    nx, ny, _ = n
    y, x = np.meshgrid((np.arange(ny) + 0.5) / ny, (np.arange(nx) + 0.5) / nx, indexing="ij")
    vc = np.zeros((5, 1, ny, nx))
    vc[0, 0] = 1.0 + (0.2 * np.cos(2 * np.pi * y) if ny > 1 else 0.0)
    vc[1, 0] = np.where(x < 0.5, -vmag, vmag)
    if ny > 1:
        vc[2, 0] = 0.5 * vmag * np.sin(2 * np.pi * y) * np.exp(-((x - 0.5) / 0.1) ** 2)
    vc[4] = prs
    return vc


def mhd_lowbeta_blast(n, p_in: float, p_out: float, b0: float = 1.0, r0: float = 0.2):
    """Blast in a strongly magnetised box (2D or 3D, n = (nx, ny, nz)): rho = 1,
    v = 0, p = p_in inside radius r0 of the centre and p_out outside, uniform
    B = b0 (1, 1, 0)/sqrt(2) on faces and cells.  With p_out << b0^2 the
    pressure recovered from the total energy goes negative behind the front
    (ConsToPrim's small_prs floor).  Returns vc[8, nz, ny, nx], vs[dims, ...]."""
    nx, ny, nz = n
    dims = 3 if nz > 1 else 2
    z, y, x = np.meshgrid((np.arange(nz) + 0.5) / nz, (np.arange(ny) + 0.5) / ny,
                          (np.arange(nx) + 0.5) / nx, indexing="ij")
    r2 = (x - 0.5) ** 2 + (y - 0.5) ** 2 + ((z - 0.5) ** 2 if dims == 3 else 0.0)
    vc = np.zeros((8, nz, ny, nx))
    vc[0] = 1.0
    vc[7] = np.where(r2 < r0 * r0, p_in, p_out)
    b = b0 / np.sqrt(2.0)
    vc[4] = b
    vc[5] = b
    vs = np.zeros((dims, nz, ny, nx))
    vs[0] = b
    vs[1] = b
    return vc, vs


def divb_max(vs: np.ndarray) -> float:
    """max |div B| (per unit volume, dx = 1/n) of a periodic face field."""
    _, nz, ny, nx = vs.shape
    d = (np.roll(vs[0], -1, axis=2) - vs[0]) * nx + (np.roll(vs[1], -1, axis=1) - vs[1]) * ny
    if nz > 1:
        d = d + (np.roll(vs[2], -1, axis=0) - vs[2]) * nz
    return float(np.max(np.abs(d)))
