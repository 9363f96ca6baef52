"""Seeded synthetic subhalo populations for the unbinding workloads.

Units follow the reference (10^10 Msun/h, Mpc/h, km/s, so G = 43.0071; reference
src/config_parser.cpp:200).  Every generator is pure numpy and deterministic for
a given seed.

Workloads (sizes as named in the benchmark definitions):
  * ``single_nfw``      one NFW halo (r_s=0.02, c=10, r_max=10 r_s), 10,000
                        particles plus 2,000 uniform interlopers inside 3 r_200
                        with 3x escape-speed velocities; m = 1e8 Msun/h.
  * ``host_with_subs``  one 2,000,000-particle NFW host with 50 embedded NFW
                        subhaloes of 500-50,000 particles (one tree over all).
  * ``population``      independent NFW subhaloes with N_p drawn from
                        dN/dN_p ~ N_p^-1.9 on [20, 1e6], log-normal
                        concentrations (median 8, 0.15 dex), 10% interlopers.
  * ``hydro``           as ``population`` with 70% DM / 20% gas (with specific
                        internal energy) / 9% stars / 1% black holes at 100x mass.
  * ``evolve``          ID-preserving evolution of a population through
                        snapshots with tidal stripping and 2% jitter.

The NFW radius is drawn by inverse CDF of the truncated profile
M(<x) ~ ln(1+x) - x/(1+x); velocities are isotropic Gaussians with the local
isotropic-Jeans dispersion of the truncated profile; the escape speed uses the
closed-form truncated-NFW potential.
"""
import os

import numpy as np

G_GRAV = 43.0071
PARTICLE_MASS = 0.01          # 1e8 Msun/h in 1e10 Msun/h
SOFTENING = 0.001             # 1 kpc/h
R200_PER_1E4 = 0.2            # r_200 of a 10,000-particle halo (r_s=0.02, c=10)

_XMAX = 64.0
_XGRID = np.concatenate([[0.0], np.geomspace(1e-7, _XMAX, 6000)])


def _mu(x):
    x = np.asarray(x, dtype=np.float64)
    return np.log1p(x) - x / (1.0 + x)


def _build_tables():
    x = _XGRID
    mu = _mu(x)
    # G(x) = int_x^XMAX mu(t) / (t^3 (1+t)^2) dt  (isotropic Jeans integrand)
    xs = x[1:]
    g = _mu(xs) / (xs ** 3 * (1.0 + xs) ** 2)
    seg = 0.5 * (g[1:] + g[:-1]) * np.diff(xs)
    tail = np.concatenate([np.cumsum(seg[::-1])[::-1], [0.0]])
    # near zero mu ~ x^2/2 so g ~ 1/(2x): extend analytically to x=0 (value unused at 0)
    gtab = np.concatenate([[tail[0] + 0.5 * np.log(xs[0] / 1e-12)], tail])
    return mu, gtab


_MU_TAB, _G_TAB = _build_tables()


def nfw_inverse_radius(u, c):
    """x = r/r_s with M(<x)/M(<c) = u for a profile truncated at x=c."""
    target = np.asarray(u) * _mu(c)
    return np.interp(target, _MU_TAB, _XGRID)


def nfw_sigma2(x, c, gm_tot, rs):
    """Isotropic Jeans 1-D velocity dispersion^2 of a truncated NFW (sigma=0 at x=c)."""
    x = np.minimum(x, c)
    gx = np.interp(x, _XGRID, _G_TAB)
    gc = np.interp(c, _XGRID, _G_TAB)
    s2 = x * (1.0 + x) ** 2 * np.maximum(gx - gc, 0.0)
    return gm_tot / (rs * _mu(c)) * s2


def nfw_potential(x, c, gm_tot, rs):
    """Closed-form potential of an NFW truncated at x=c (total mass M_tot)."""
    x = np.asarray(x, dtype=np.float64)
    inside = gm_tot / (rs * _mu(c)) * (_mu(np.maximum(x, 1e-12)) / np.maximum(x, 1e-12)
                                         + 1.0 / (1.0 + x) - 1.0 / (1.0 + c))
    outside = gm_tot / (rs * np.maximum(x, 1e-12))
    return -np.where(x < c, inside, outside)


def nfw_vmax2(c, gm_tot, rs):
    xm = 2.1626
    return gm_tot / (rs * _mu(c)) * _mu(np.minimum(xm, c)) / np.minimum(xm, c)


def _unit_vectors(rng, n):
    z = rng.uniform(-1.0, 1.0, n)
    phi = rng.uniform(0.0, 2.0 * np.pi, n)
    s = np.sqrt(np.maximum(0.0, 1.0 - z * z))
    return np.stack([s * np.cos(phi), s * np.sin(phi), z], axis=1)


class Population:
    """Particles grouped by subhalo: SoA float32 arrays plus CSR offsets."""

    def __init__(self, pos, vel, mass, offsets, pid, ptype=None, u=None,
                 ref_pos=None, ref_vel=None, meta=None):
        self.pos = np.ascontiguousarray(pos, dtype=np.float32)
        self.vel = np.ascontiguousarray(vel, dtype=np.float32)
        self.mass = np.ascontiguousarray(mass, dtype=np.float32)
        self.offsets = np.ascontiguousarray(offsets, dtype=np.int64)
        self.pid = np.ascontiguousarray(pid, dtype=np.int64)
        self.ptype = None if ptype is None else np.ascontiguousarray(ptype, dtype=np.int32)
        self.u = None if u is None else np.ascontiguousarray(u, dtype=np.float32)
        self.meta = meta or {}
        if ref_pos is None or ref_vel is None:
            ref_pos, ref_vel = mass_weighted_frame(self.pos, self.vel, self.mass, self.offsets)
        self.ref_pos = np.ascontiguousarray(ref_pos, dtype=np.float32)
        self.ref_vel = np.ascontiguousarray(ref_vel, dtype=np.float32)

    @property
    def n(self):
        return int(self.offsets[-1])

    @property
    def nsub(self):
        return len(self.offsets) - 1

    def counts(self):
        return np.diff(self.offsets)

    def subset(self, sub_ids):
        """A new Population holding only the listed subhaloes (in that order)."""
        sub_ids = np.asarray(sub_ids, dtype=np.int64)
        cnt = self.counts()[sub_ids]
        idx = np.concatenate([np.arange(self.offsets[s], self.offsets[s + 1]) for s in sub_ids]) \
            if len(sub_ids) else np.zeros(0, np.int64)
        off = np.concatenate([[0], np.cumsum(cnt)])
        return Population(self.pos[idx], self.vel[idx], self.mass[idx], off, self.pid[idx],
                          None if self.ptype is None else self.ptype[idx],
                          None if self.u is None else self.u[idx],
                          self.ref_pos[sub_ids], self.ref_vel[sub_ids],
                          dict(self.meta, parent_subset=True))


def mass_weighted_frame(pos, vel, mass, offsets):
    """Mass-weighted mean position/velocity of every subhalo (float64 sums)."""
    nsub = len(offsets) - 1
    nz = np.diff(offsets) > 0
    starts = np.asarray(offsets[:-1])[nz]
    m = mass.astype(np.float64)
    msum = np.ones(nsub)
    rp = np.zeros((nsub, 3))
    rv = np.zeros((nsub, 3))
    if len(starts):
        m32 = np.ascontiguousarray(mass, dtype=np.float32)[:, None]
        msum[nz] = np.add.reduceat(m, starts)
        rp[nz] = np.add.reduceat(pos * m32, starts, axis=0, dtype=np.float64)
        rv[nz] = np.add.reduceat(vel * m32, starts, axis=0, dtype=np.float64)
    return (rp / msum[:, None]).astype(np.float32), (rv / msum[:, None]).astype(np.float32)


def _nfw_members(rng, n, c, rs, gm_tot):
    x = nfw_inverse_radius(rng.uniform(0.0, 1.0, n), c)
    r = x * rs
    dirs = _unit_vectors(rng, n)
    sig = np.sqrt(nfw_sigma2(x, c, gm_tot, rs))
    v = rng.normal(0.0, 1.0, (n, 3)) * sig[:, None]
    return dirs * r[:, None], v


def _interlopers(rng, n, c, rs, gm_tot, rmax):
    r = rmax * np.cbrt(rng.uniform(0.0, 1.0, n))
    dirs = _unit_vectors(rng, n)
    vesc = np.sqrt(2.0 * np.abs(nfw_potential(r / rs, c, gm_tot, rs)))
    v = rng.normal(0.0, 1.0, (n, 3)) * (np.sqrt(3.0) * vesc)[:, None]
    return dirs * r[:, None], v


def single_nfw(seed=0, n_halo=10000, n_inter=2000, rs=0.02, c=10.0, centre=(1.0, 1.0, 1.0)):
    """One NFW halo with uniform interlopers (the CPU-runnable single-subhalo case)."""
    rng = np.random.default_rng(seed)
    m = PARTICLE_MASS
    gm = G_GRAV * n_halo * m
    x1, v1 = _nfw_members(rng, n_halo, c, rs, gm)
    x2, v2 = _interlopers(rng, n_inter, c, rs, gm, 3.0 * c * rs)
    pos = np.concatenate([x1, x2]) + np.asarray(centre)
    vel = np.concatenate([v1, v2])
    n = n_halo + n_inter
    return Population(pos, vel, np.full(n, m), [0, n], np.arange(n),
                      meta=dict(kind="single_nfw", seed=seed, n_halo=n_halo, n_inter=n_inter,
                                rs=rs, c=c))


def draw_counts(rng, nsub, nmin=20, nmax=1_000_000, slope=1.9):
    a = 1.0 - slope
    u = rng.uniform(0.0, 1.0, nsub)
    n = (nmin ** a + u * (nmax ** a - nmin ** a)) ** (1.0 / a)
    return np.clip(np.floor(n).astype(np.int64), nmin, nmax)


def population(seed=2, nsub=20000, target_particles=None, nmin=20, nmax=1_000_000,
               slope=1.9, interloper_frac=0.10, box=50.0, hydro=False, chunk=4_000_000):
    """Independent NFW subhaloes with power-law particle counts.

    If ``target_particles`` is given, subhaloes are drawn until the total
    (members + interlopers) reaches it; otherwise exactly ``nsub`` are drawn.
    """
    rng = np.random.default_rng(seed)
    if target_particles is not None:
        counts = []
        total = 0
        while total < target_particles:
            c = draw_counts(rng, 4096, nmin, nmax, slope)
            tot = c + np.floor(interloper_frac * c).astype(np.int64)
            cs = np.cumsum(tot) + total
            k = int(np.searchsorted(cs, target_particles)) + 1
            counts.append(c[:k])
            total = int(cs[min(k, len(cs)) - 1])
        nmem = np.concatenate(counts)
    else:
        nmem = draw_counts(rng, nsub, nmin, nmax, slope)
    nsub = len(nmem)
    ninter = np.floor(interloper_frac * nmem).astype(np.int64)
    ntot = nmem + ninter
    conc = 10.0 ** rng.normal(np.log10(8.0), 0.15, nsub)
    r200 = R200_PER_1E4 * np.cbrt(nmem / 1e4)
    rs = r200 / conc
    centres = rng.uniform(0.1 * box, 0.9 * box, (nsub, 3))
    bulk = rng.normal(0.0, 100.0, (nsub, 3))
    offsets = np.concatenate([[0], np.cumsum(ntot)]).astype(np.int64)
    N = int(offsets[-1])
    pos = np.empty((N, 3), np.float32)
    vel = np.empty((N, 3), np.float32)
    mass = np.full(N, PARTICLE_MASS, np.float32)
    ptype = np.empty(N, np.int32) if hydro else None
    uint = np.zeros(N, np.float32) if hydro else None
    # chunked over subhaloes to bound memory
    s0 = 0
    while s0 < nsub:
        s1 = s0 + 1
        while s1 < nsub and offsets[s1 + 1] - offsets[s0] <= chunk:
            s1 += 1
        subs = np.arange(s0, s1)
        _fill_population_chunk(rng, subs, nmem, ninter, conc, rs, centres, bulk, offsets,
                               pos, vel, mass, ptype, uint, hydro)
        s0 = s1
    pid = np.arange(N, dtype=np.int64)
    return Population(pos, vel, mass, offsets, pid, ptype, uint,
                      meta=dict(kind="hydro" if hydro else "population", seed=seed, nsub=nsub,
                                nmem=nmem, ninter=ninter, conc=conc, rs=rs, centres=centres,
                                bulk=bulk, r200=r200))


def _fill_population_chunk(rng, subs, nmem, ninter, conc, rs, centres, bulk, offsets,
                           pos, vel, mass, ptype, uint, hydro):
    nm = nmem[subs]
    ni = ninter[subs]
    # member particles
    sid_m = np.repeat(subs, nm)
    c_m = conc[sid_m]
    rs_m = rs[sid_m]
    gm_m = G_GRAV * PARTICLE_MASS * nmem[sid_m]
    x = nfw_inverse_radius(rng.uniform(0.0, 1.0, len(sid_m)), c_m)
    dirs = _unit_vectors(rng, len(sid_m))
    sig = np.sqrt(nfw_sigma2(x, c_m, gm_m, rs_m))
    xm = dirs * (x * rs_m)[:, None]
    vm = rng.normal(0.0, 1.0, (len(sid_m), 3)) * sig[:, None]
    # interlopers: uniform inside 3 r_200 with per-component sigma = sqrt(3) v_esc
    sid_i = np.repeat(subs, ni)
    c_i = conc[sid_i]
    rs_i = rs[sid_i]
    gm_i = G_GRAV * PARTICLE_MASS * nmem[sid_i]
    ri = 3.0 * c_i * rs_i * np.cbrt(rng.uniform(0.0, 1.0, len(sid_i)))
    vesc = np.sqrt(2.0 * np.abs(nfw_potential(ri / rs_i, c_i, gm_i, rs_i)))
    xi = _unit_vectors(rng, len(sid_i)) * ri[:, None]
    vi = rng.normal(0.0, 1.0, (len(sid_i), 3)) * (np.sqrt(3.0) * vesc)[:, None]
    # scatter into place: members first, then interlopers, per subhalo
    start = offsets[subs]
    rank_m = np.arange(len(sid_m)) - np.repeat(np.cumsum(nm) - nm, nm)
    dst_m = np.repeat(start, nm) + rank_m
    rank_i = np.arange(len(sid_i)) - np.repeat(np.cumsum(ni) - ni, ni)
    dst_i = np.repeat(start + nm, ni) + rank_i
    pos[dst_m] = xm + centres[sid_m]
    vel[dst_m] = vm + bulk[sid_m]
    pos[dst_i] = xi + centres[sid_i]
    vel[dst_i] = vi + bulk[sid_i]
    if hydro:
        n_all = len(dst_m) + len(dst_i)
        dst = np.concatenate([dst_m, dst_i])
        sid = np.concatenate([sid_m, sid_i])
        draw = rng.uniform(0.0, 1.0, n_all)
        t = np.where(draw < 0.70, 1, np.where(draw < 0.90, 0, np.where(draw < 0.99, 4, 5)))
        ptype[dst] = t
        # gas: u log-normal (0.2 dex) around 0.3 v_circ^2 with v_circ = V_max
        gmx = G_GRAV * PARTICLE_MASS * nmem[sid]
        vmax2 = nfw_vmax2(conc[sid], gmx, rs[sid])
        ugas = 0.3 * vmax2 * 10.0 ** rng.normal(0.0, 0.2, n_all)
        uint[dst] = np.where(t == 0, ugas, 0.0)
        # black holes: 100x mass, uniform within 0.1 r_s of the centre, local Jeans velocity
        bh = t == 5
        if bh.any():
            d = dst[bh]
            s = sid[bh]
            rb = 0.1 * rs[s] * np.cbrt(rng.uniform(0.0, 1.0, len(d)))
            pos[d] = _unit_vectors(rng, len(d)) * rb[:, None] + centres[s]
            sgb = np.sqrt(nfw_sigma2(rb / rs[s], conc[s], G_GRAV * PARTICLE_MASS * nmem[s], rs[s]))
            vel[d] = rng.normal(0.0, 1.0, (len(d), 3)) * sgb[:, None] + bulk[s]
            mass[d] = 100.0 * PARTICLE_MASS


def host_with_subs(seed=1, n_host=2_000_000, nsub=50, nsub_min=500, nsub_max=50_000, c=10.0):
    """One NFW host (self-similar r_200) with embedded NFW subhaloes; one particle set."""
    rng = np.random.default_rng(seed)
    m = PARTICLE_MASS
    r200 = R200_PER_1E4 * np.cbrt(n_host / 1e4)
    rs = r200 / c
    gm = G_GRAV * n_host * m
    xh, vh = _nfw_members(rng, n_host, c, rs, gm)
    sig_host = np.sqrt(np.mean(vh ** 2))
    ns = np.floor(np.exp(rng.uniform(np.log(nsub_min), np.log(nsub_max), nsub))).astype(np.int64)
    parts_x = [xh]
    parts_v = [vh]
    for k in range(nsub):
        r200s = R200_PER_1E4 * np.cbrt(ns[k] / 1e4)
        rss = r200s / c
        xs, vs = _nfw_members(rng, int(ns[k]), c, rss, G_GRAV * ns[k] * m)
        cen = _unit_vectors(rng, 1)[0] * 0.5 * r200 * np.cbrt(rng.uniform())
        parts_x.append(xs + cen)
        parts_v.append(vs + rng.normal(0.0, sig_host, 3))
    pos = np.concatenate(parts_x) + np.array([25.0, 25.0, 25.0])
    vel = np.concatenate(parts_v)
    n = len(pos)
    return Population(pos, vel, np.full(n, m), [0, n], np.arange(n),
                      meta=dict(kind="host_with_subs", seed=seed, n_host=n_host, sub_counts=ns))


class Hierarchy:
    """Nested subhaloes for Subhalo_t::RecursiveUnbind (src/subhalo_unbind.cpp:810-840):
    pop holds every subhalo's own particles (one subhalo per node), children[s]
    its NestedSubhalos, roots the top of every tree, track / sink the TrackId and
    SinkTrackId (-1 when not trapped) of every subhalo."""

    def __init__(self, pop, children, roots, track, sink):
        self.pop = pop
        self.children = children
        self.roots = roots
        self.track = track
        self.sink = sink


def hierarchy(seed=5, nhost=3, n_host=20000, kids=(2, 4), n_kid=(500, 5000), n_grandkid=(100, 1500),
              p_overlap=0.35):
    """nhost trees: an NFW host (10% interlopers) with embedded NFW children, each
    with 0-2 embedded grandchildren and, under the first child of every host, one
    more level.  A share p_overlap of the grandchildren sit on their parent's
    centre at its bulk velocity, so the phase-space merger test fires; the last
    host's first child is already trapped (SinkTrackId set)."""
    rng = np.random.default_rng(seed)
    m = PARTICLE_MASS
    xs, vs, counts, children, roots = [], [], [], [], []

    def add(n, c, centre, bulk, inter=0):
        r200 = R200_PER_1E4 * np.cbrt(n / 1e4)
        rs = r200 / c
        gm = G_GRAV * n * m
        x, v = _nfw_members(rng, n, c, rs, gm)
        if inter:
            xi, vi = _interlopers(rng, inter, c, rs, gm, 3.0 * r200)
            x, v = np.concatenate([x, xi]), np.concatenate([v, vi])
        xs.append(x + centre)
        vs.append(v + bulk)
        counts.append(len(x))
        children.append([])
        return len(counts) - 1, r200, float(np.sqrt(np.mean(v[:n] ** 2)))

    def embed(parent, cen, bulk, r200, sig, nrange, overlap):
        n = int(np.exp(rng.uniform(np.log(nrange[0]), np.log(nrange[1]))))
        if overlap:
            c2, b2 = cen.copy(), bulk.copy()
        else:
            c2 = cen + _unit_vectors(rng, 1)[0] * 0.5 * r200 * np.cbrt(rng.uniform(0.2, 1.0))
            b2 = bulk + rng.normal(0.0, sig, 3)
        k, r, sg = add(n, 8.0, c2, b2)
        children[parent].append(k)
        return k, c2, b2, r, sg

    for h in range(nhost):
        cen = np.array([10.0 + 10.0 * h, 25.0, 25.0])
        bulk = rng.normal(0.0, 50.0, 3)
        root, r200, sig = add(n_host, 10.0, cen, bulk, inter=n_host // 10)
        roots.append(root)
        for j in range(int(rng.integers(kids[0], kids[1] + 1))):
            kid, kc, kb, kr, ks = embed(root, cen, bulk, r200, sig, n_kid, False)
            for _ in range(int(rng.integers(0, 3))):
                embed(kid, kc, kb, kr, ks, n_grandkid, rng.uniform() < p_overlap)
            if j == 0:
                g, gc, gb, gr, gs = embed(kid, kc, kb, kr, ks, n_grandkid, False)
                embed(g, gc, gb, gr, gs, (60, 200), False)
    pos = np.concatenate(xs)
    vel = np.concatenate(vs)
    off = np.concatenate([[0], np.cumsum(counts)])
    n = int(off[-1])
    pop = Population(pos, vel, np.full(n, m), off, np.arange(n), meta=dict(kind="hierarchy", seed=seed))
    ns = len(counts)
    track = np.arange(ns, dtype=np.int64)
    sink = np.full(ns, -1, np.int64)
    sink[children[roots[-1]][0]] = roots[-1]
    return Hierarchy(pop, children, roots, track, sink)


_NOISE_STREAMS = 16  # fixed, so the noise does not depend on the thread count


def _jitter_fill(gens, pool, out, jitter):
    """out = 1 + jitter * N(0,1), filled by independent streams over fixed chunks."""
    chunks = np.array_split(out.reshape(-1), _NOISE_STREAMS)

    def fill(k):
        c = chunks[k]
        gens[k].standard_normal(out=c, dtype=np.float32)
        c *= np.float32(jitter)
        c += np.float32(1.0)

    list(pool.map(fill, range(_NOISE_STREAMS)))


def evolve_iter(pop, nsnap=8, seed=4, shrink=0.9, kick=1.5, jitter=0.02, dt=0.002):
    """ID-preserving snapshot sequence of a population (a generator).

    Snapshot k strips particles outside a tidal radius r_200 * shrink**k by a
    velocity kick of ``kick`` x the local escape speed (in a random direction),
    moves kicked particles ballistically, and applies ``jitter`` relative
    Gaussian noise to every particle's offset from its subhalo centre and
    velocity.  Particles keep their ids and their slot in the arrays.
    """
    from concurrent.futures import ThreadPoolExecutor
    rng = np.random.default_rng(seed)
    gens = [np.random.default_rng(s) for s in np.random.SeedSequence(seed).spawn(_NOISE_STREAMS)]
    pool = ThreadPoolExecutor(min(_NOISE_STREAMS, os.cpu_count() or 1))
    meta = pop.meta
    nsub = pop.nsub
    counts = pop.counts()
    sid = np.repeat(np.arange(nsub), counts)
    centre = np.repeat(meta["centres"].astype(np.float32), counts, axis=0)
    bulk = np.repeat(meta["bulk"].astype(np.float32), counts, axis=0)
    rs = meta["rs"]
    conc = meta["conc"]
    gm = G_GRAV * PARTICLE_MASS * meta["nmem"]
    r200 = (conc * rs).astype(np.float32)[sid]
    dx = pop.pos - centre
    dv = pop.vel - bulk
    kicked = np.zeros(pop.n, bool)
    kdt = np.zeros((pop.n, 1), np.float32)  # dt for kicked particles, else 0
    noise = np.empty_like(dx)
    for k in range(nsnap):
        if k > 0:
            _jitter_fill(gens, pool, noise, jitter)
            dx *= noise
            _jitter_fill(gens, pool, noise, jitter)
            dv *= noise
            np.multiply(dv, kdt, out=noise)
            dx += noise
        r = np.sqrt(np.einsum("ij,ij->i", dx, dx))
        newk = np.where((~kicked) & (r > r200 * np.float32(shrink ** k)))[0]
        if len(newk):
            s = sid[newk]
            vesc = np.sqrt(2.0 * np.abs(nfw_potential(r[newk] / rs[s], conc[s], gm[s], rs[s])))
            dv[newk] += (_unit_vectors(rng, len(newk)) * (kick * vesc)[:, None]).astype(np.float32)
            kicked[newk] = True
            kdt[newk] = np.float32(dt)
        yield Population(dx + centre, dv + bulk, pop.mass, pop.offsets, pop.pid, pop.ptype, pop.u,
                         meta=dict(meta, snapshot=k, kicked=kicked.copy()))


def evolve(pop, nsnap=8, seed=4, shrink=0.9, kick=1.5, jitter=0.02, dt=0.002):
    """evolve_iter as a list."""
    return list(evolve_iter(pop, nsnap, seed, shrink, kick, jitter, dt))


class HostHalos:
    """Host haloes for particle reassignment: particles host halo by host halo
    (group_offsets), each halo's subhaloes contiguous inside it, a tracer flag and
    the subhalo of every particle."""

    def __init__(self, pos, is_tracer, sub, group_offsets, sub_offsets):
        self.pos, self.is_tracer, self.sub = pos, is_tracer, sub
        self.group_offsets, self.sub_offsets = group_offsets, sub_offsets

    @property
    def n(self):
        return len(self.sub)


def host_haloes(seed=7, ngroups=40, subs=(1, 8), n_sub=(50, 3000), tracer_frac=0.6, stray=0.2, rs=0.02,
                box=None, spread=0.05):
    """Host haloes of a hydrodynamical run: every subhalo an NFW clump around a
    centre near its host's, a fraction tracer_frac of its particles tracers (dark
    matter and stars), and a fraction `stray` of its gas drawn from a neighbouring
    subhalo's clump (the particles that reassignment moves).  With box set, host
    centres are uniform in the box and positions wrapped into it."""
    rng = np.random.default_rng(seed)
    pos_l, tr_l, sub_l, goff, soff = [], [], [], [0], [0]
    sid = 0
    for g in range(ngroups):
        hc = rng.uniform(0, box, 3) if box else rng.uniform(0.5, 1.5, 3) + 3.0 * g
        ns = int(rng.integers(subs[0], subs[1] + 1))
        centres = hc + rng.normal(0, spread, (ns, 3)) * (np.arange(ns) > 0)[:, None]
        sizes = rng.integers(n_sub[0], n_sub[1] + 1, ns)
        scale = rs * (sizes / sizes.max()) ** (1 / 3) + 1e-3
        for s in range(ns):
            m = int(sizes[s])
            r = scale[s] * nfw_inverse_radius(rng.random(m), 10.0)
            p = centres[s] + r[:, None] * _unit_vectors(rng, m)
            tr = rng.random(m) < tracer_frac
            gas = np.nonzero(~tr)[0]
            if ns > 1 and len(gas):
                move = gas[rng.random(len(gas)) < stray]
                other = (s + rng.integers(1, ns, len(move))) % ns
                rr = scale[other] * nfw_inverse_radius(rng.random(len(move)), 10.0)
                p[move] = centres[other] + rr[:, None] * _unit_vectors(rng, len(move))
            pos_l.append(p)
            tr_l.append(tr)
            sub_l.append(np.full(m, sid, np.int32))
            sid += 1
            soff.append(soff[-1] + m)
        goff.append(soff[-1])
    pos = np.concatenate(pos_l)
    if box:
        pos = np.mod(pos, box)
    return HostHalos(pos.astype(np.float32), np.concatenate(tr_l).astype(np.uint8), np.concatenate(sub_l),
                     np.asarray(goff, np.int64), np.asarray(soff, np.int64))
