"""Seeded synthetic particle boxes in Abacus's slab layout.

A box of ``cpd**3`` cells in the periodic unit box [-0.5, 0.5)^3 (grid.cpp:30-58,
cpd odd).  Slab ``x`` holds cells (x, y, z) in y-major, z-minor order
(cellparticles.cpp:30-36); its particles are stored PosXYZ-style as one float32
array ``X[n] | Y[n] | Z[n]`` (cellparticles.cpp:218-228) with positions
relative to the cell centre (grid.cpp:98-107, 174-178), computed in float64
and rounded once to float32.

Workloads:
  * ``uniform_box``   N particles uniform in the box from a seed; the cell
    counts are one multinomial draw, the same law as binning N iid points.
  * ``clustered_box`` 70% lattice points displaced by a Gaussian random field
    with P(k) ~ k^-2 (rms displacement 0.5 cell) plus 30% in Plummer spheres
    whose masses follow dN/dM ~ M^-1.9, scale radii 0.02-0.3 cell widths.
  * ``clustered_slabs`` the same recipe at production size (CPD 255, ~1e9
    particles), realised only in the slabs a run needs.
"""
import os

import numpy as np

CELLINFO_DTYPE = np.dtype(
    [("startindex", "<u8"), ("startindex_with_ghost", "<u8"), ("count", "<i4"), ("active", "<i4"),
     ("mean_square_velocity", "<f4"), ("max_component_velocity", "<f4"),
     ("max_component_acceleration", "<f4")], align=True)
assert CELLINFO_DTYPE.itemsize == 40


class Slab:
    """One slab: PosXYZ positions and per-cell counts/starts."""

    def __init__(self, pos, count):
        self.count = np.ascontiguousarray(count, dtype=np.int32)
        self.start = np.concatenate([[0], np.cumsum(self.count, dtype=np.int64)[:-1]]).astype(np.int64)
        self.n = int(self.count.sum())
        self.pos = np.ascontiguousarray(pos, dtype=np.float32)
        assert self.pos.shape == (3 * self.n,)
        ci = np.zeros(self.count.size, CELLINFO_DTYPE)
        ci["startindex"] = self.start
        ci["startindex_with_ghost"] = self.start
        ci["count"] = self.count
        ci["active"] = self.count
        self.cellinfo = ci

    @property
    def xyz(self):
        """(n, 3) view-copy of the positions."""
        return self.pos.reshape(3, self.n).T


class Box:
    def __init__(self, cpd, slabs, name=""):
        assert cpd % 2 == 1, "cpd must be odd (grid.cpp:50)"
        self.cpd = cpd
        self.slabs = slabs
        self.name = name
        self.N = sum(s.n for s in slabs)

    def slab(self, x):
        return self.slabs[x % self.cpd]

    def counts(self):
        return np.stack([s.count for s in self.slabs])  # [x, y*cpd+z]

    def mean_spacing(self):
        return 1.0 / np.cbrt(self.N)


def cell_centre(c, cpd):
    """grid::CellCenter, grid.cpp:98-107 (one axis)."""
    return (np.asarray(c, dtype=np.float64) - (cpd - 1) // 2) / cpd


def box_from_global(pos, cpd, name=""):
    """Bin global float64 positions into cells and build the slab layout."""
    pos = np.asarray(pos, dtype=np.float64)
    pos = pos - np.floor(pos + 0.5)                         # wrap to [-0.5, 0.5)
    cell = np.floor((pos + 0.5) * cpd).astype(np.int64)     # Position2Cell, grid.cpp:133-137
    cell = np.clip(cell, 0, cpd - 1)
    local = (pos - cell_centre(cell, cpd)).astype(np.float32)
    # Bin by slab (a radix sort on the small slab index), then by cell within
    # each slab; stable, so particles keep their input order inside a cell.
    sx = cell[:, 0].astype(np.uint16)
    order = np.argsort(sx, kind="stable")
    bounds = np.concatenate([[0], np.cumsum(np.bincount(sx, minlength=cpd))])
    slabs = []
    for x in range(cpd):
        idx = order[bounds[x]:bounds[x + 1]]
        cyz = cell[idx, 1] * cpd + cell[idx, 2]
        o2 = np.argsort(cyz, kind="stable")
        xyz = local[idx[o2]]
        cnt = np.bincount(cyz, minlength=cpd * cpd).astype(np.int32)
        slabs.append(Slab(np.concatenate([xyz[:, 0], xyz[:, 1], xyz[:, 2]]), cnt))
    return Box(cpd, slabs, name)


def uniform_box(N, cpd, seed):
    """N particles uniform in the periodic box, generated slab by slab."""
    rng = np.random.default_rng(seed)
    ncell = cpd ** 3
    counts = rng.multinomial(N, np.full(ncell, 1.0 / ncell)).reshape(cpd, cpd * cpd)
    slabs = []
    for x in range(cpd):
        cnt = counts[x]
        n = int(cnt.sum())
        u = rng.random((n, 3))
        cyz = np.repeat(np.arange(cpd * cpd), cnt)
        ci = np.stack([np.full(n, x), cyz // cpd, cyz % cpd], axis=1)
        glob = (ci + u) / cpd - 0.5
        local = (glob - cell_centre(ci, cpd)).astype(np.float32)
        slabs.append(Slab(np.concatenate([local[:, 0], local[:, 1], local[:, 2]]), cnt))
    return Box(cpd, slabs, f"uniform N={N} cpd={cpd} seed={seed}")


def _grf_displacement(ng, rms, rng):
    """Displacement field of a Gaussian random density field with P(k) ~ k^-2
    on an ng^3 lattice (Zel'dovich: psi_k = i k/k^2 delta_k), scaled so the
    rms displacement per axis component is ``rms`` (box units)."""
    k1 = np.fft.fftfreq(ng) * ng
    kz1 = np.fft.rfftfreq(ng) * ng
    kx, ky, kz = np.meshgrid(k1, k1, kz1, indexing="ij")
    k2 = kx * kx + ky * ky + kz * kz
    k2[0, 0, 0] = 1.0
    amp = np.sqrt(k2 ** -1.0)                      # sqrt(P(k)), P ~ k^-2
    amp[0, 0, 0] = 0.0
    noise = rng.standard_normal((ng, ng, ng))
    dk = np.fft.rfftn(noise) * amp
    out = np.empty((ng ** 3, 3), np.float64)
    for a, ka in enumerate((kx, ky, kz)):
        psi = np.fft.irfftn(1j * ka / k2 * dk, s=(ng, ng, ng), axes=(0, 1, 2))
        out[:, a] = psi.ravel()
    out *= rms / np.sqrt((out ** 2).mean(axis=0).mean())
    return out


def clustered_box(n1d, cpd, seed, halo_frac=0.3, n_halos=2000, rmin=0.02, rmax=0.3,
                  mass_slope=1.9, mass_range=4.0, rtrunc=10.0):
    """The clustered recipe: N = n1d^3; 70% lattice + GRF displacement, 30% in
    Plummer spheres.  The lattice fraction is a uniform random 70% subset of
    the n1d^3 lattice.  Halo masses are drawn from dN/dM ~ M^-1.9 on
    [M0, mass_range * M0] and normalised to the halo particle budget, so the
    mean halo holds halo_frac*N/n_halos particles (2,517 for 256^3) and the
    largest about 2.8x that: cell occupancies then run from 0 to ~5,000.
    Scale radii are log-uniform in [rmin, rmax] cell widths; each sphere is
    truncated at ``rtrunc`` scale radii."""
    rng = np.random.default_rng(seed)
    N = n1d ** 3
    n_halo_parts = int(round(halo_frac * N))
    n_lat = N - n_halo_parts
    w = 1.0 / cpd
    # lattice + GRF displacements
    disp = _grf_displacement(n1d, 0.5 * w, rng)
    keep = np.sort(rng.choice(N, size=n_lat, replace=False))
    idx = np.stack(np.unravel_index(keep, (n1d, n1d, n1d)), axis=1)
    lat = (idx + 0.5) / n1d - 0.5 + disp[keep]
    # Plummer spheres
    a = 1.0 - mass_slope
    lo, hi = 1.0, float(mass_range)
    u = rng.random(n_halos)
    m = (lo ** a + u * (hi ** a - lo ** a)) ** (1.0 / a)
    nh = np.floor(m / m.sum() * n_halo_parts).astype(np.int64)
    short = n_halo_parts - nh.sum()
    nh[np.argsort(-m)[:short]] += 1
    centres = rng.random((n_halos, 3)) - 0.5
    radii = np.exp(rng.uniform(np.log(rmin), np.log(rmax), n_halos)) * w
    hc = np.repeat(np.arange(n_halos), nh)
    # Plummer: enclosed mass fraction X = r^3/(r^2+a^2)^1.5 -> r = a / sqrt(X^-2/3 - 1),
    # truncated at rtrunc * a
    xmax = rtrunc ** 3 / (rtrunc ** 2 + 1.0) ** 1.5
    X = rng.random(hc.size) * xmax
    r = radii[hc] / np.sqrt(X ** (-2.0 / 3.0) - 1.0)
    v = rng.standard_normal((hc.size, 3))
    v /= np.linalg.norm(v, axis=1)[:, None]
    halo = centres[hc] + r[:, None] * v
    pos = np.concatenate([lat, halo])
    return box_from_global(pos, cpd, f"clustered n1d={n1d} cpd={cpd} seed={seed}")


def _grf_mesh(ng, rms, rng):
    """_grf_displacement on an ng^3 mesh, as float32 (ng, ng, ng, 3)."""
    return _grf_displacement(ng, rms, rng).astype(np.float32).reshape(ng, ng, ng, 3)


def _plane_displacement(mesh, x, n1d):
    """Trilinear interpolation of the mesh on the whole lattice plane at global
    x: the plane's (n1d, n1d, 3) displacements (lattice points at
    ((j, k) + 0.5)/n1d - 0.5 in y and z)."""
    ng = mesh.shape[0]

    def axis_weights(q):
        u = (np.asarray(q, np.float64) + 0.5) * ng - 0.5
        i0 = np.floor(u).astype(np.int64)
        f = (u - i0).astype(np.float32)
        return i0 % ng, (i0 + 1) % ng, f

    ix0, ix1, fx = axis_weights(x)
    sl = (1.0 - fx) * mesh[ix0] + fx * mesh[ix1]                        # (ng, ng, 3) over (y, z)
    q = (np.arange(n1d) + 0.5) / n1d - 0.5
    i0, i1, f = axis_weights(q)
    sl = (1.0 - f)[:, None, None] * sl[i0] + f[:, None, None] * sl[i1]    # (n1d, ng, 3)
    return (1.0 - f)[None, :, None] * sl[:, i0] + f[None, :, None] * sl[:, i1]


def _trilinear(mesh, q):
    """Periodic trilinear interpolation of a cell-centred mesh (value (i,j,k)
    at ((i,j,k) + 0.5)/ng - 0.5) at global positions q (n, 3)."""
    ng = mesh.shape[0]
    u = (q + 0.5) * ng - 0.5
    i0 = np.floor(u).astype(np.int64)
    f = (u - i0).astype(np.float32)
    out = np.zeros((q.shape[0], 3), np.float32)
    for dx in (0, 1):
        wx = f[:, 0] if dx else 1.0 - f[:, 0]
        ix = (i0[:, 0] + dx) % ng
        for dy in (0, 1):
            wy = f[:, 1] if dy else 1.0 - f[:, 1]
            iy = (i0[:, 1] + dy) % ng
            for dz in (0, 1):
                wz = f[:, 2] if dz else 1.0 - f[:, 2]
                iz = (i0[:, 2] + dz) % ng
                out += (wx * wy * wz)[:, None] * mesh[ix, iy, iz]
    return out


def _in_slabs(x_lo, x_hi, cpd, wanted):
    """Whether the x interval [x_lo, x_hi] (global, may leave the box) touches
    any slab in the boolean array ``wanted``."""
    c0 = int(np.floor((x_lo + 0.5) * cpd))
    c1 = int(np.floor((x_hi + 0.5) * cpd))
    if c1 - c0 + 1 >= cpd:
        return bool(wanted.any())
    return bool(wanted[np.arange(c0, c1 + 1) % cpd].any())


def _in_slabs_many(x_lo, x_hi, cpd, wanted):
    """_in_slabs over arrays of intervals, each at most 2 cpd wide and within
    one box length of the box."""
    c0 = np.floor((np.asarray(x_lo) + 0.5) * cpd).astype(np.int64)
    c1 = np.floor((np.asarray(x_hi) + 0.5) * cpd).astype(np.int64)
    cum = np.concatenate([[0], np.cumsum(np.tile(wanted, 4))])     # wanted counts over [-cpd, 3 cpd)
    n = cum[np.clip(c1, -cpd, 3 * cpd - 1) + cpd + 1] - cum[np.clip(c0, -cpd, 3 * cpd - 1) + cpd]
    return np.where(c1 - c0 + 1 >= cpd, bool(wanted.any()), n > 0)


def clustered_slabs(n1d, cpd, seed, slabs, grf_mesh=256, halos_per_cell=2000 / 69 ** 3, halo_frac=0.3,
                    rmin=0.02, rmax=0.3, mass_slope=1.9, mass_range=4.0, rtrunc=10.0, lattice_frac=0.7,
                    mesh=None):
    """The clustered recipe of ``clustered_box`` at any size, realised only in
    the listed slabs (every other slab of the returned Box is empty).

    Every part of the realisation has its own seed, so the particles of a slab
    do not depend on which other slabs are asked for:
      * lattice plane i (x index) keeps each of its n1d^2 points with
        probability ``lattice_frac`` (seed (seed, 1, i)) and displaces them by
        a P(k) ~ k^-2 Gaussian random field of rms 0.5 cell, computed on a
        ``grf_mesh``^3 mesh (seed (seed, 0)) and interpolated trilinearly;
      * ``halos_per_cell * cpd^3`` Plummer spheres (2,000 at CPD 69) hold
        ``halo_frac`` of n1d^3 particles; their masses (dN/dM ~ M^-1.9 on
        [M0, mass_range M0]), centres and scale radii (log-uniform in
        [rmin, rmax] cells) come from seed (seed, 2), and halo h's particles
        from seed (seed, 3, h).
    The nominal particle count (for the mean spacing) is n1d^3.  ``mesh``
    passes in the GRF mesh of these arguments when the caller already has it."""
    w = 1.0 / cpd
    wanted = np.zeros(cpd, bool)
    wanted[np.asarray(list(slabs)) % cpd] = True
    if mesh is None:
        mesh = _grf_mesh(grf_mesh, 0.5 * w, np.random.default_rng([seed, 0]))
    maxd = float(np.abs(mesh).max())
    parts = []

    def wanted_only(p):
        # wrapped into the box; only the particles of wanted slabs are kept
        p = p - np.floor(p + 0.5)
        sx = np.clip(np.floor((p[:, 0] + 0.5) * cpd).astype(np.int64), 0, cpd - 1)
        return p[wanted[sx]]

    ax = (np.arange(n1d) + 0.5) / n1d - 0.5
    gy, gz = np.meshgrid(ax, ax, indexing="ij")
    gy, gz = gy.ravel(), gz.ravel()
    for i in range(n1d):
        if not _in_slabs(ax[i] - maxd, ax[i] + maxd, cpd, wanted):
            continue
        rng = np.random.default_rng([seed, 1, i])
        keep = rng.random(n1d * n1d) < lattice_frac
        d = _plane_displacement(mesh, ax[i], n1d).reshape(-1, 3)[keep]
        parts.append(wanted_only(np.stack([ax[i] + d[:, 0], gy[keep] + d[:, 1], gz[keep] + d[:, 2]], axis=1)))
    rng = np.random.default_rng([seed, 2])
    n_halos = max(1, int(round(halos_per_cell * cpd ** 3)))
    n_halo_parts = int(round(halo_frac * n1d ** 3))
    a = 1.0 - mass_slope
    u = rng.random(n_halos)
    m = (1.0 + u * (float(mass_range) ** a - 1.0)) ** (1.0 / a)
    nh = np.floor(m / m.sum() * n_halo_parts).astype(np.int64)
    nh[np.argsort(-m)[:n_halo_parts - int(nh.sum())]] += 1
    centres = rng.random((n_halos, 3)) - 0.5
    radii = np.exp(rng.uniform(np.log(rmin), np.log(rmax), n_halos)) * w
    xmax = rtrunc ** 3 / (rtrunc ** 2 + 1.0) ** 1.5
    ext = rtrunc * radii
    near = _in_slabs_many(centres[:, 0] - ext, centres[:, 0] + ext, cpd, wanted) & (nh > 0)
    for h in np.flatnonzero(near):
        hr = np.random.default_rng([seed, 3, h])
        X = hr.random(nh[h]) * xmax
        r = radii[h] / np.sqrt(X ** (-2.0 / 3.0) - 1.0)
        v = hr.standard_normal((nh[h], 3))
        v /= np.linalg.norm(v, axis=1)[:, None]
        parts.append(wanted_only(centres[h] + r[:, None] * v))
    pos = np.concatenate(parts) if parts else np.zeros((0, 3))
    box = box_from_global(pos, cpd, f"clustered n1d={n1d} cpd={cpd} seed={seed} slabs {int(wanted.sum())}/{cpd}")
    box.N = n1d ** 3
    return box


_PAR_MESH = None  # the GRF mesh a forked generator pool inherits


def _clustered_chunk(job):
    n1d, cpd, seed, chunk, kw = job
    box = clustered_slabs(n1d, cpd, seed, chunk, mesh=_PAR_MESH, **kw)
    return [(x % cpd, box.slab(x)) for x in chunk]


def clustered_slabs_parallel(n1d, cpd, seed, slabs, workers=None, chunk=4, progress=None, **kw):
    """``clustered_slabs`` realised by a pool of forked processes, ``chunk``
    slabs per job: the same particles slab for slab (each slab's particles
    depend only on the seed), in a fraction of the time at production size
    (CPD 255, ~1e9 particles).  The GRF mesh is computed once, here, and
    inherited by the workers.  Forks: call it before the process touches a
    GPU.  progress(done, total), if given, is called as jobs complete."""
    import multiprocessing as mp
    global _PAR_MESH
    want = sorted({x % cpd for x in slabs})
    grf = kw.pop("grf_mesh", 256)
    _PAR_MESH = _grf_mesh(grf, 0.5 / cpd, np.random.default_rng([seed, 0]))
    kw["grf_mesh"] = grf
    jobs = [(n1d, cpd, seed, want[i:i + chunk], kw) for i in range(0, len(want), chunk)]
    workers = workers or min(len(jobs), 16, os.cpu_count() or 1)
    out = [Slab(np.zeros(0, np.float32), np.zeros(cpd * cpd, np.int32)) for _ in range(cpd)]
    try:
        with mp.get_context("fork").Pool(workers) as pool:
            for done, part in enumerate(pool.imap_unordered(_clustered_chunk, jobs), 1):
                for x, sl in part:
                    out[x] = sl
                if progress:
                    progress(done, len(jobs))
    finally:
        _PAR_MESH = None
    box = Box(cpd, out, f"clustered n1d={n1d} cpd={cpd} seed={seed} slabs {len(want)}/{cpd}")
    box.N = n1d ** 3
    return box


# --- softening / FOF parameters -------------------------------------------

SPLINE_FACTOR = 2.15517   # single spline internal length / Plummer length (singlestep.cpp:116-118)


def eps_param(softening, plummer_len):
    """The kernel's eps argument, NearFieldDriver ctor (directdriver.cpp:105-117),
    from the Plummer-equivalent softening length (singlestep.cpp:113-125)."""
    if softening == "single_spline":
        e = SPLINE_FACTOR * plummer_len
        return np.float32(1.0 / (e * e)), e
    if softening == "plummer":
        e = plummer_len
        return np.float32(e * e), e
    if softening == "cubic_spline":
        e = 1.10064 * plummer_len
        return np.float32(1.0 / e), e
    if softening == "cubic_plummer":
        e = plummer_len
        return np.float32(e ** 3), e
    raise ValueError(softening)


def fof_b2(mean_spacing, linking_length=0.25):
    """DensityKernelRad2 (proepi.cpp:686-694): (b * spacing)^2 * (1 + 1e-5)."""
    b = linking_length * mean_spacing
    return np.float32(b * b * (1.0 + 1.0e-5))


SOFTENINGS = {"plummer": 0, "single_spline": 1, "cubic_spline": 2, "cubic_plummer": 3}
