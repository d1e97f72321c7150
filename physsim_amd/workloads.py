"""Seeded synthetic inputs for the benchmark and parity workloads.

All generators use numpy's Philox4x32-10 bit generator with seed
`config_id * 1000 + replicate` and round every value to FP32, so the FP32
device path and the FP64 oracle see bit-identical inputs.  Charges are integer
multiples of 0.1 (main.js:511); spins are surface-velocity fractions turned
into angular celerity as in main.js:518-521.
"""
import numpy as np

from . import _abi as A


def rng(config_id, replicate=0):
    return np.random.Generator(np.random.Philox(config_id * 1000 + replicate))


def f32(a):
    return np.asarray(a, dtype=np.float32).astype(np.float64)


def spin_to_angw(sv, mass):
    """main.js:517-521: clamp to +-0.999, angw = sign*|sv| / (r*sqrt(1 - sv^2)), r = cbrt(m)."""
    sv = np.clip(sv, -0.999, 0.999)
    r = np.cbrt(mass)
    a = np.abs(sv)
    return np.where(a > 0, np.sign(sv) * a / (r * np.sqrt(1 - a * a)), 0.0)


def disk(n, radius, cx, cy, soft, g, rel_noise=0.05):
    """Uniform disk r = R*sqrt(U), lognormal(0, 0.5) masses, charges 0.1*k with
    k in [-10, 10], circular speed from the enclosed mass plus N(0, 0.05) noise."""
    r = radius * np.sqrt(g.random(n))
    phi = 2 * np.pi * g.random(n)
    x = cx + r * np.cos(phi)
    y = cy + r * np.sin(phi)
    m = g.lognormal(0.0, 0.5, n)
    k = g.integers(-10, 11, n)
    q = k * 0.1
    order = np.argsort(r)
    menc = np.empty(n)
    menc[order] = np.cumsum(m[order]) - m[order]
    vc = np.sqrt(np.maximum(menc * r * r / (r * r + soft * soft) ** 1.5, 0.0))
    vc = np.minimum(vc, 0.5)
    wx = -vc * np.sin(phi) + g.normal(0, rel_noise, n)
    wy = vc * np.cos(phi) + g.normal(0, rel_noise, n)
    return dict(x=f32(x), y=f32(y), wx=f32(wx), wy=f32(wy), mass=f32(m), charge=f32(q), angw=np.zeros(n))


def config1(n=1024, replicate=0):
    """Plain method, small: gravity + Coulomb, softening 0.01, theta 0.5, dt 1e-3,
    despawn boundary; disk of radius 10 at the centre of a 64x64 domain."""
    g = rng(1, replicate)
    st = disk(n, 10.0, 32.0, 32.0, 0.01, g)
    prm = dict(toggles0=A.GRAVITY | A.COULOMB | A.BARNES_HUT, boundary_mode=A.BOUND_DESPAWN,
               domain_w=64.0, domain_h=64.0, softening=0.01, theta=0.5)
    return st, prm, 1e-3


def config2(n=65536, replicate=0, side=None, open_box=False):
    """Relativistic torus box (min-image), gravity + Coulomb + Yukawa (mu = 1) +
    magnetic dipole, spins uniform in [-1, 1]; softening 8, theta 0.5, dt 1/128.
    The box side scales with sqrt(n) to keep the density of 65536 in 256x256.
    Leaf sources at the current time (signal delay off): the reference would
    take them from the light cone, and from a history that needs 128 updates
    before any leaf acts at all (forces.js:652-655, integrator.js:1549-1565).
    open_box: the same particles and forces in an open (despawn) square, where
    Barnes-Hut converges to the direct sum."""
    g = rng(2, replicate)
    L = side if side is not None else 256.0 * np.sqrt(n / 65536.0)
    x = g.random(n) * L
    y = g.random(n) * L
    m = g.lognormal(0.0, 0.5, n)
    q = g.integers(-10, 11, n) * 0.1
    wx = g.normal(0, 0.05, n)
    wy = g.normal(0, 0.05, n)
    sv = g.uniform(-1.0, 1.0, n)
    angw = spin_to_angw(sv, f32(m))
    st = dict(x=f32(x), y=f32(y), wx=f32(wx), wy=f32(wy), mass=f32(m), charge=f32(q), angw=f32(angw))
    prm = dict(toggles0=A.GRAVITY | A.COULOMB | A.YUKAWA | A.MAGNETIC | A.RELATIVITY | A.BARNES_HUT,
               boundary_mode=A.BOUND_LOOP, topology=A.TORUS, periodic_mode=A.PERIODIC_MIN_IMAGE,
               domain_w=float(L), domain_h=float(L), softening=8.0, theta=0.5, yukawa_mu=1.0,
               signal_delay=A.DELAY_OFF)
    if open_box:
        prm["boundary_mode"] = A.BOUND_DESPAWN
    return st, prm, 1.0 / 128


def gaussian_field(G, g, rms, index=-2.0):
    """Gaussian random field with power spectrum P(k) ~ k^index, zero mean, given rms."""
    white = np.fft.fft2(g.normal(size=(G, G)))
    k = np.sqrt(np.fft.fftfreq(G)[:, None] ** 2 + np.fft.fftfreq(G)[None, :] ** 2)
    k[0, 0] = np.inf
    f = np.fft.ifft2(white * k ** (index / 2)).real
    f -= f.mean()
    return f * (rms / f.std())


def config3(n=1 << 20, replicate=0, grid=128, side=1024.0, total_mass=16.0):
    """Projected Plummer sphere (scale radius 20, truncated at 0.45 W) with
    virialised Gaussian velocities; gravity with the 1PN Einstein-Infeld-Hoffmann
    terms (which the reference evaluates only with gravitomagnetism on,
    forces.js:242-248, so that is on too), Higgs and axion fields on a grid^2
    PQS grid with FFT self-gravity; initial fields Gaussian random with
    spectral index -2 and rms 0.01 about the vacuum (Higgs 1, axion 0).
    Masses: lognormal(0, 0.5) scaled to total_mass, so the virial speed stays
    well below c.  Relativity is off, so with the reference's coupling of
    signal delay to relativity the 1PN velocity-Verlet pass takes no leaf
    sources (they would come from a light cone nothing has recorded,
    forces.js:542-546): only accepted nodes."""
    g = rng(3, replicate)
    a = 20.0
    rmax = 0.45 * side
    u = g.random(n) * (rmax ** 2 / (rmax ** 2 + a ** 2))
    R = a * np.sqrt(u / (1 - u))
    phi = 2 * np.pi * g.random(n)
    x = side / 2 + R * np.cos(phi)
    y = side / 2 + R * np.sin(phi)
    m = g.lognormal(0.0, 0.5, n)
    m *= total_mass / m.sum()
    # virial ratio K/|W| = 1/2 with the Plummer binding energy |W| = (3 pi / 32) M^2 / a
    sigma = np.sqrt((3 * np.pi / 32) * total_mass / (2 * a))
    wx = g.normal(0, sigma, n)
    wy = g.normal(0, sigma, n)
    st = dict(x=f32(x), y=f32(y), wx=f32(wx), wy=f32(wy), mass=f32(m), charge=np.zeros(n), angw=np.zeros(n))
    prm = dict(toggles0=A.GRAVITY | A.GRAVITOMAG | A.ONE_PN | A.HIGGS | A.AXION | A.BARNES_HUT,
               boundary_mode=A.BOUND_DESPAWN, domain_w=float(side), domain_h=float(side), softening=8.0, theta=0.5,
               field_grid=grid)
    fields = dict(higgs=f32(1.0 + gaussian_field(grid, g, 0.01)), axion=f32(gaussian_field(grid, g, 0.01)))
    return st, prm, 1.0 / 128, fields


def make_params(lib_default, **kw):
    for k, v in kw.items():
        setattr(lib_default, k, v)
    return lib_default


def plummer_charged(n=65536, replicate=0, a=20.0, side=512.0):
    """Training systems of the learned far-field correction: a projected
    Plummer sphere of scale radius a (truncated at 0.45 side), lognormal(0, 0.5)
    masses, charges 0.1*k for k uniform in [-10, 10], at rest; gravity and
    Coulomb with the Barnes-Hut softening of the N=65536 box workload."""
    g = rng(5, replicate)
    rmax = 0.45 * side
    u = g.random(n) * (rmax ** 2 / (rmax ** 2 + a ** 2))
    R = a * np.sqrt(u / (1 - u))
    phi = 2 * np.pi * g.random(n)
    x = side / 2 + R * np.cos(phi)
    y = side / 2 + R * np.sin(phi)
    m = g.lognormal(0.0, 0.5, n)
    q = g.integers(-10, 11, n) * 0.1
    z = np.zeros(n)
    st = dict(x=f32(x), y=f32(y), wx=z, wy=z.copy(), mass=f32(m), charge=f32(q), angw=z.copy())
    prm = dict(toggles0=A.GRAVITY | A.COULOMB | A.BARNES_HUT, boundary_mode=A.BOUND_DESPAWN,
               domain_w=float(side), domain_h=float(side), softening=8.0, theta=0.3)
    return st, prm


def config4(n_systems=512, replicate=0, n_min=256, n_max=16384):
    """Batched ensemble: n_systems independent config-1 disks (gravity +
    Coulomb, softening 0.01, theta 0.5, dt 1e-3, despawn) with N_s = floor(n_min
    (n_max / n_min)^U), log-uniform in [n_min, n_max].  Returns the per-system
    states, the shared params and dt."""
    g = rng(4, replicate)
    u = g.random(n_systems)
    sizes = np.minimum(np.floor(n_min * (n_max / n_min) ** u).astype(int), n_max)
    states = []
    prm = None
    for s, n in enumerate(sizes):
        st, prm, dt = config1(int(n), replicate=1000 * (replicate + 1) + s)
        states.append(st)
    return states, prm, dt
