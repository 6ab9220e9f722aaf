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
"""
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
    cid = (cell[:, 0] * cpd + cell[:, 1]) * cpd + cell[:, 2]
    order = np.argsort(cid, kind="stable")
    cid = cid[order]
    local = local[order]
    counts = np.bincount(cid, minlength=cpd ** 3).reshape(cpd, cpd * cpd)
    slabs = []
    off = 0
    for x in range(cpd):
        n = int(counts[x].sum())
        xyz = local[off:off + n]
        slabs.append(Slab(np.concatenate([xyz[:, 0], xyz[:, 1], xyz[:, 2]]), counts[x]))
        off += n
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
