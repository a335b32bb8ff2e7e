"""Synthetic GALIC models: parameters, force tables, targets and particles.

The reference builds these at start-up (main.c:51-80): parameters from a file,
structural scalars (structure.c:54-143), grid geometry and tables
(grid.c:329-891), Jeans dispersions (disp_fields.c:502-1049) and particles
(set_particles.c:15-307).  Only the hot path is in scope here, so this module
prepares its inputs directly from the configurations' physical numbers:

* positions follow the reference's inverse-CDF samplers (halo.c:17-40,
  disk.c:17-51, bulge.c:15-36) driven by a counter-based RNG (rng.py);
* Tint follows set_particles.c:87, Vesc grid.c:115-120 with the same field;
* the density target is vol*rho at cell centres, smoothed, with the fine level
  renormalised to the component mass (grid.c:830-877);
* the dispersion targets are CIC maps of sampled points (grid.c:941-1076);
* dispersions come from spherical Jeans integrals (halo, bulge) and a
  vertically isothermal estimate (disk): set-up inputs, not hot-path code;
* a disk's tabulated field is the Miyamoto-Nagai potential with the disk's
  scale length and height; galic_tree_fill produces the tree version.

Everything returned is plain numpy, handed to the C ABI unchanged.
"""
import ctypes
from dataclasses import dataclass, field

import numpy as np

from . import _abi
from .rng import gaussians, uniforms

# internal units: kpc, 1e10 Msun, km/s (init.c:40-63 with allvars.h:227,245)
UNIT_LENGTH = 3.085678e21
UNIT_MASS = 1.989e43
UNIT_VELOCITY = 1e5
GRAVITY = 6.6738e-8
HUBBLE = 3.2407789e-18
UNIT_TIME = UNIT_LENGTH / UNIT_VELOCITY
G_INTERNAL = GRAVITY / UNIT_LENGTH ** 3 * UNIT_MASS * UNIT_TIME ** 2
H_INTERNAL = HUBBLE * UNIT_TIME
MIN_DENSITY = 1e-25  # allvars.h:201


@dataclass
class Config:
    """A model configuration.  Masses in 1e10 Msun, lengths in kpc."""
    name: str = "halo"
    n: int = 100000
    halo_mass: float = 100.0
    halo_a: float = 30.0
    halo_beta: float = 0.0
    halo_tovs: int = 0
    disk_mass: float = 0.0
    disk_h: float = 3.5
    disk_z0: float = 0.35
    disk_frac: float = 0.0       # fraction of particles in the disk
    disk_tovs: int = 3
    bulge_mass: float = 0.0
    bulge_a: float = 0.7
    bulge_frac: float = 0.0
    bulge_tovs: int = 0
    outer_frac: float = (30.0 / 31.0) ** 2   # truncation at 30 a (grid.c:336)
    inner_frac: float = 1e-7
    fg_nbin: int = 128
    level: int = 6
    torbit_fac: float = 4.0
    timestep_orbit: float = 0.01
    timestep_cell: float = 0.25
    min_per_bin_density: int = 50
    min_per_bin_disp: int = 100
    max_vel_vesc: float = 0.9999
    halo_in_table: bool = False  # SampleForceNhalo > 0: tabulate the exact halo field
    seed: int = 1
    target_samples: int = 0      # points per component for the dispersion targets (0: auto)
    softening: float = 0.05
    theta: float = 0.4
    max_steps: int = 100000000


CONFIGS = {
    # Plain method, small: single Hernquist halo, N=1e5, 128^2 table, 64^2 grids
    "halo-1e5": Config(name="halo-1e5", n=100000),
    # Precision form: beta=0.3 halo + exponential sech^2 disk, N=1e6, 256^2 table
    "halo-disk-1e6": Config(name="halo-disk-1e6", n=1000000, halo_beta=0.3, halo_tovs=1, disk_mass=4.0,
                            disk_frac=0.2, fg_nbin=256),
    # Batching form: halo + disk + bulge, N=4e6
    "composite-4e6": Config(name="composite-4e6", n=4000000, halo_beta=0.3, halo_tovs=1, disk_mass=4.0,
                            disk_frac=0.2, bulge_mass=1.0, bulge_frac=0.05, fg_nbin=256),
    # Full v1.1 model: isotropic halo and bulge, disk of velocity structure 3 with sigma_R = sigma_z,
    # N=2e7, 512^2 force tables, 512^2 response grids (level 9)
    "composite-2e7": Config(name="composite-2e7", n=20000000, halo_tovs=0, disk_mass=4.0, disk_frac=0.2,
                            disk_tovs=3, bulge_mass=1.0, bulge_frac=0.05, bulge_tovs=0, fg_nbin=512, level=9),
}


def config(name, **overrides):
    base = CONFIGS[name]
    d = dict(base.__dict__)
    d.update(overrides)
    return Config(**d)


@dataclass
class Workload:
    cfg: Config
    params: _abi.GalicParams
    pot: np.ndarray
    dpotdr: np.ndarray
    dpotdz: np.ndarray
    vstream: dict                     # type -> FG_Nbin^2 table
    targets: dict                     # type -> list of 6 stacks (Tm, MT, Er, Et, Ep, Eq)
    pos: np.ndarray                   # (n,3)
    vel: np.ndarray                   # (n,3)
    mass: np.ndarray
    tint: np.ndarray
    ptype: np.ndarray                 # int32
    vesc: np.ndarray
    disp: np.ndarray                  # (n,4) target vr2, vt2, vp2, vq2 (set_particles.c:137-161)
    vstr: np.ndarray                  # streaming velocity at each particle
    extra: dict = field(default_factory=dict)

    @property
    def n(self):
        return len(self.mass)

    @property
    def types(self):
        return [t for t in (1, 2, 3) if self.params.NType[t] > 0]


# --------------------------------------------------------------- structure --

def make_params(cfg):
    """Structural scalars (structure.c:54-143) for directly given masses and radii."""
    p = _abi.GalicParams()
    p.G = G_INTERNAL
    m200 = cfg.halo_mass + cfg.disk_mass + cfg.bulge_mass
    p.M200 = m200
    p.V200 = (m200 * 10 * G_INTERNAL * H_INTERNAL) ** (1.0 / 3.0)
    p.R200 = p.V200 / (10 * H_INTERNAL)
    p.Halo_Mass = cfg.halo_mass
    p.Halo_A = cfg.halo_a
    p.Disk_Mass = cfg.disk_mass
    p.Disk_H = cfg.disk_h
    p.Disk_Z0 = cfg.disk_z0
    p.Bulge_Mass = cfg.bulge_mass
    p.Bulge_A = cfg.bulge_a
    nd = int(round(cfg.n * cfg.disk_frac)) if cfg.disk_mass > 0 else 0
    nb = int(round(cfg.n * cfg.bulge_frac)) if cfg.bulge_mass > 0 else 0
    nh = cfg.n - nd - nb
    p.MType[1], p.MType[2], p.MType[3] = cfg.halo_mass, cfg.disk_mass, cfg.bulge_mass
    p.NType[1], p.NType[2], p.NType[3] = nh, nd, nb
    p.SampleForceNhalo = 1 if cfg.halo_in_table else 0
    p.SampleForceNdisk = 1 if cfg.disk_mass > 0 else 0
    p.SampleForceNbulge = 0
    p.TypeOfVelocityStructure[1] = cfg.halo_tovs
    p.TypeOfVelocityStructure[2] = cfg.disk_tovs
    p.TypeOfVelocityStructure[3] = cfg.bulge_tovs
    p.TimeStepFactorOrbit = cfg.timestep_orbit
    p.TimeStepFactorCellCross = cfg.timestep_cell
    p.MaxVelInUnitsVesc = cfg.max_vel_vesc
    p.MaxStepsPerOrbit = cfg.max_steps
    p.LowerDispLimit = (0.01 * p.V200) ** 2  # structure.c:65
    p.MinParticlesPerBinForDensityMeasurement = cfg.min_per_bin_density
    p.MinParticlesPerBinForDispersionMeasurement = cfg.min_per_bin_disp
    for t in range(4):
        p.fac_value_rs[t] = p.fac_value_ts[t] = p.fac_value_qs[t] = p.fac_value_ps[t] = 1.0
    p.FG_Nbin = cfg.fg_nbin
    p.DG_MaxLevel = cfg.level
    p.EG_MaxLevel = cfg.level
    p.ErrTolTheta = cfg.theta
    p.ForceSoftening = 2.8 * cfg.softening  # init.c:78-84
    _abi.check(_abi.lib().galic_grid_create(ctypes.byref(p), cfg.outer_frac, cfg.inner_frac))
    return p


def stack_offset(l, i, j):
    return ((1 << (2 * l)) - 1) // 3 + i * (1 << l) + j


def nstack(L):
    return ((1 << (2 + 2 * L)) - 1) // 3


def smooth_stack(data, L):
    """Quadtree sums of the fine level into the coarse levels (grid.c:1247-1265)."""
    for l in range(L - 1, -1, -1):
        n = 1 << l
        fine = data[stack_offset(l + 1, 0, 0): stack_offset(l + 1, 0, 0) + 4 * n * n].reshape(2 * n, 2 * n)
        s = fine[0::2, 0::2] + fine[1::2, 0::2] + fine[0::2, 1::2] + fine[1::2, 1::2]
        data[stack_offset(l, 0, 0): stack_offset(l, 0, 0) + n * n] = s.ravel()
    return data


# ------------------------------------------------------------ analytic fields --

def halo_density(p, x, y, z):
    r = np.sqrt(x * x + y * y + z * z)
    rho = p.Halo_Mass / (2 * np.pi) * p.Halo_A / (r + 1.0e-6 * p.Halo_A) / (r + p.Halo_A) ** 3
    return np.where(np.abs(rho) < MIN_DENSITY, 0.0, rho)


def disk_density(p, x, y, z):
    if p.Disk_Mass <= 0:
        return np.zeros_like(x)
    R = np.sqrt(x * x + y * y)
    e = np.exp(-np.abs(z) / p.Disk_Z0)
    sech = 2 * e / (1 + e * e)   # 2/(exp(z/z0)+exp(-z/z0)) without overflow
    rho = p.Disk_Mass / (4 * np.pi * p.Disk_H * p.Disk_H * p.Disk_Z0) * np.exp(-R / p.Disk_H) * sech ** 2
    return np.where(np.abs(rho) < MIN_DENSITY, 0.0, rho)


def bulge_density(p, x, y, z):
    # bulge.c:39-48 multiplies by Bulge_Mass where Bulge_A belongs; kept as is
    if p.Bulge_Mass <= 0:
        return np.zeros_like(x)
    r = np.sqrt(x * x + y * y + z * z)
    rho = p.Bulge_Mass / (2 * np.pi) * p.Bulge_Mass / (r + 1.0e-6 * p.Bulge_A) / (r + p.Bulge_A) ** 3
    return np.where(np.abs(rho) < MIN_DENSITY, 0.0, rho)


DENSITY = {1: halo_density, 2: disk_density, 3: bulge_density}


def hernquist_accfac(G, M, A, r):
    return G * M / ((r + 1.0e-6 * A) * (r + A) * (r + A))


def mn_disk_field(p, R, z):
    """Miyamoto-Nagai potential and accelerations (radial, vertical) for the disk."""
    s = np.sqrt(z * z + p.Disk_Z0 ** 2)
    D = np.sqrt(R * R + (p.Disk_H + s) ** 2)
    pot = -p.G * p.Disk_Mass / D
    aR = -p.G * p.Disk_Mass * R / D ** 3
    az = -p.G * p.Disk_Mass * z * (p.Disk_H + s) / (s * D ** 3)
    return pot, aR, az


def fg_nodes(p):
    j = np.arange(p.FG_Nbin)
    return p.FG_Rmin * (np.power(p.FG_Fac, j) - 1.0)


def make_tables(p, cfg):
    """FG_Pot, FG_DPotDR, FG_DPotDz at nodes R_j = Rmin(Fac^j-1), z_k likewise (grid.c:766-786)."""
    nodes = fg_nodes(p)
    R = nodes[None, :].repeat(p.FG_Nbin, 0)   # [k][j]: z row k, R column j
    z = nodes[:, None].repeat(p.FG_Nbin, 1)
    pot = np.zeros_like(R)
    aR = np.zeros_like(R)
    az = np.zeros_like(R)
    if p.Disk_Mass > 0:
        dp, dR, dz = mn_disk_field(p, R, z)
        pot += dp
        aR += dR
        az += dz
    if p.SampleForceNhalo > 0:
        r = np.sqrt(R * R + z * z)
        pot += -p.G * p.Halo_Mass / (r + p.Halo_A)
        f = hernquist_accfac(p.G, p.Halo_Mass, p.Halo_A, r)
        aR += -f * R
        az += -f * z
    return np.ascontiguousarray(pot.ravel()), np.ascontiguousarray(aR.ravel()), np.ascontiguousarray(az.ravel())


def model_field(p, x, y, z):
    """Acceleration and potential of the model (exact forms of what the tables hold)."""
    R = np.sqrt(x * x + y * y)
    r = np.sqrt(x * x + y * y + z * z)
    f = hernquist_accfac(p.G, p.Halo_Mass, p.Halo_A, r)
    ax, ay, az = -f * x, -f * y, -f * z
    pot = -p.G * p.Halo_Mass / (r + p.Halo_A)
    if p.Disk_Mass > 0:
        dp, dR, dz = mn_disk_field(p, R, np.abs(z))
        Rs = np.where(R > 0, R, 1.0)
        ax = ax + dR * np.where(R > 0, x / Rs, 1.0)
        ay = ay + dR * np.where(R > 0, y / Rs, 0.0)
        az = az + np.sign(z) * dz
        pot = pot + dp
    if p.Bulge_Mass > 0:
        fb = hernquist_accfac(p.G, p.Bulge_Mass, p.Bulge_A, r)
        ax, ay, az = ax - fb * x, ay - fb * y, az - fb * z
        pot = pot - p.G * p.Bulge_Mass / (r + p.Bulge_A)
    return np.stack([ax, ay, az], -1), pot


# ------------------------------------------------------------------ sampling --

def sample_halo(p, seed, stream, ids, a=None):
    """halo.c:17-40 (spherical halo): r = a(q+sqrt q)/(1-q), isotropic, r <= Rmax."""
    a = p.Halo_A if a is None else a
    out = np.zeros((len(ids), 3))
    todo = np.arange(len(ids))
    attempt = 0
    while len(todo):
        u = uniforms(seed, stream, ids[todo], attempt)
        q = u[0]
        r = a * (q + np.sqrt(q)) / (1 - q)
        phi = u[1] * np.pi * 2
        theta = np.arccos(u[2] * 2 - 1)
        x = np.stack([r * np.sin(theta) * np.cos(phi), r * np.sin(theta) * np.sin(phi), r * np.cos(theta)], -1)
        ok = np.sqrt((x * x).sum(-1)) <= p.Rmax
        out[todo[ok]] = x[ok]
        todo = todo[~ok]
        attempt += 1
    return out


def sample_disk(p, seed, stream, ids):
    """disk.c:17-51: sech^2 height, exponential radius by Newton iteration."""
    out = np.zeros((len(ids), 3))
    todo = np.arange(len(ids))
    attempt = 0
    while len(todo):
        u = uniforms(seed, stream, ids[todo], attempt)
        z = p.Disk_Z0 / 2 * np.log(u[0] / (1 - u[0]))
        q = u[1]
        R = np.ones_like(q)
        for _ in range(200):
            f = (1 + R) * np.exp(-R) + q - 1
            fp = -R * np.exp(-R)
            Rn = R - f / fp
            done = np.abs(Rn - R) / np.abs(Rn) <= 1e-7
            R = Rn
            if done.all():
                break
        R *= p.Disk_H
        phi = u[2] * np.pi * 2
        x = np.stack([R * np.cos(phi), R * np.sin(phi), z], -1)
        ok = (x * x).sum(-1) <= p.Rmax * p.Rmax
        out[todo[ok]] = x[ok]
        todo = todo[~ok]
        attempt += 1
    return out


def sample_type(p, t, seed, stream, ids):
    if t == 1:
        return sample_halo(p, seed, stream, ids)
    if t == 2:
        return sample_disk(p, seed, stream, ids)
    return sample_halo(p, seed, stream, ids, a=p.Bulge_A)


# ------------------------------------------------------------- dispersions --

def _enclosed_mass(p, r):
    m = p.Halo_Mass * (r / (r + p.Halo_A)) ** 2
    if p.Disk_Mass > 0:
        m = m + p.Disk_Mass * (1 - (1 + r / p.Disk_H) * np.exp(-r / p.Disk_H))
    if p.Bulge_Mass > 0:
        m = m + p.Bulge_Mass * (r / (r + p.Bulge_A)) ** 2
    return m


def spherical_jeans(p, t, beta, rgrid):
    """sigma_r^2(r) = r^(-2 beta)/rho int_r^inf rho r'^(2 beta) G M(r')/r'^2 dr' on a log grid."""
    A = p.Halo_A if t == 1 else p.Bulge_A
    rr = np.geomspace(rgrid.min() * 0.5, max(rgrid.max(), p.Rmax) * 1e3, 8000)
    rho = 1.0 / (rr + 1e-6 * A) / (rr + A) ** 3  # shape only; normalisation cancels
    integrand = rho * rr ** (2 * beta) * p.G * _enclosed_mass(p, rr) / rr ** 2
    lr = np.log(rr)
    seg = 0.5 * (integrand[1:] * rr[1:] + integrand[:-1] * rr[:-1]) * np.diff(lr)
    tail = np.concatenate([np.cumsum(seg[::-1])[::-1], [0.0]])
    s2 = tail / (rho * rr ** (2 * beta))
    return np.interp(np.log(rgrid), lr, s2)


def hernquist_sigma2_isotropic(p, r):
    """Closed form of the isotropic Hernquist dispersion (halo.c:95-118)."""
    m, a = p.Halo_Mass, p.Halo_A
    x = r / a
    return p.G * m / (12.0 * a) * np.abs(12 * r * (r + a) ** 3 / a ** 4 * np.log((r + a) / r) -
                                         r / (r + a) * (25 + x * (52 + 42 * x + 12 * x * x)))


def disk_moments(p, R):
    """Vertically isothermal sech^2 disk: sigma_z^2 = pi G Sigma z0; sigma_R = sigma_z;
    streaming from the Jeans equation with k = 1 (streaming = <vphi^2> - sigma_R^2)."""
    sigma = p.Disk_Mass / (2 * np.pi * p.Disk_H ** 2) * np.exp(-R / p.Disk_H)
    s2 = np.pi * p.G * sigma * p.Disk_Z0
    acc, _ = model_field(p, R, np.zeros_like(R), np.zeros_like(R))
    vc2 = R * np.abs(acc[..., 0])
    vphi2 = vc2 + s2 * (1 - 2 * R / p.Disk_H)
    vstr2 = np.maximum(vphi2 - s2, 0.0)
    return s2, np.sqrt(vstr2)


def disp_at(p, t, tovs, beta, x, y, z, rgrid_cache):
    """Targets disp_r, disp_t, disp_p, disp_q of get_disp_rtp (disp_fields.c:66-115) and vstream."""
    r = np.sqrt(x * x + y * y + z * z)
    R = np.sqrt(x * x + y * y)
    if t == 2:
        s2, vstr = disk_moments(p, R)
        disp_t = s2
        disp_r = s2
        disp_p = s2
        disp_q = disp_p + vstr * vstr
        return np.stack([disp_r, disp_t, disp_p, disp_q], -1), vstr
    key = (t, beta)
    if key not in rgrid_cache:
        rg = np.geomspace(p.FG_Rmin * 0.1, p.Rmax * 2, 4000)
        if t == 1 and beta == 0 and p.Disk_Mass == 0 and p.Bulge_Mass == 0:
            s2 = hernquist_sigma2_isotropic(p, rg)
        else:
            s2 = spherical_jeans(p, t, beta, rg)
        rgrid_cache[key] = (np.log(rg), s2)
    lr, s2g = rgrid_cache[key]
    disp_r = np.interp(np.log(np.maximum(r, 1e-30)), lr, s2g)
    disp_t = (1 - beta) * disp_r if tovs == 1 else disp_r
    return np.stack([disp_r, disp_t, disp_t, disp_t], -1), np.zeros_like(r)


# ---------------------------------------------------------------- targets --

def density_target(p, t):
    """DGs_MassTarget: vol*rho at cell centres, smoothed, fine level renormalised (grid.c:830-877)."""
    L, nb = p.DG_MaxLevel, 1 << p.DG_MaxLevel
    j = np.arange(nb)
    r1 = p.DG_Rmin * (np.power(p.DG_Fac, j) - 1.0)
    r2 = p.DG_Rmin * (np.power(p.DG_Fac, j + 1) - 1.0)
    R1, Z1 = np.meshgrid(r1, r1)   # [k][j]
    R2, Z2 = np.meshgrid(r2, r2)
    vol = np.pi * (R2 * R2 - R1 * R1) * (Z2 - Z1) * 2
    rho = DENSITY[t](p, 0.5 * (R1 + R2), np.zeros_like(R1), 0.5 * (Z1 + Z2))
    st = np.zeros(nstack(L))
    off = stack_offset(L, 0, 0)
    st[off:] = (vol * rho).ravel()
    smooth_stack(st, L)
    msum = st[off:].sum()
    if msum > 0:
        st[off:] *= p.MType[t] / msum   # coarse levels keep the pre-renormalisation sums
    return st


def energy_targets(p, cfg, t, nsamples, seed):
    """EGs_MassTarget and EGs_EgyTarget_{r,t,p,q}: CIC of sampled points (grid.c:941-1076)."""
    L, nb = p.EG_MaxLevel, 1 << p.EG_MaxLevel
    ids = np.arange(nsamples, dtype=np.uint64)
    pos = sample_type(p, t, seed, 100 + t, ids)
    x, y, z = pos[:, 0], pos[:, 1], pos[:, 2]
    keep = np.sqrt(x * x + z * z) < p.Rmax   # the reference tests x and z only (grid.c:985)
    x, y, z = x[keep], y[keep], z[keep]
    tovs = p.TypeOfVelocityStructure[t]
    beta = cfg.halo_beta if t == 1 else 0.0
    disp, _ = disp_at(p, t, tovs, beta, x, y, z, {})
    iR, iz, fR, fz = cell_np(x, y, z, p.EG_Rmin, p.EG_Fac, nb)
    w = [(1 - fR) * (1 - fz), fR * (1 - fz), (1 - fR) * fz, fR * fz]
    cells = [iz * nb + iR, iz * nb + iR + 1, (iz + 1) * nb + iR, (iz + 1) * nb + iR + 1]
    maps = np.zeros((5, nb * nb))
    vals = [np.ones_like(x), disp[:, 0], disp[:, 1], disp[:, 2], disp[:, 3]]
    for k in range(4):
        for f in range(5):
            maps[f] += np.bincount(cells[k], weights=w[k] * vals[f], minlength=nb * nb)
    maps *= p.MType[t] / float(nsamples)
    out = []
    off = stack_offset(L, 0, 0)
    for f in range(5):
        st = np.zeros(nstack(L))
        st[off:] = maps[f]
        smooth_stack(st, L)
        out.append(st)
    return out  # MT, Er, Et, Ep, Eq


def cell_np(x, y, z, Rmin, Fac, nb):
    """Cell-centred log index (grid.c:494-534)."""
    r = np.sqrt(x * x + y * y)
    br = np.log(r / Rmin + 1.0) / np.log(Fac) - 0.5
    bz = np.log(np.abs(z) / Rmin + 1.0) / np.log(Fac) - 0.5
    br = np.maximum(br, 0)
    bz = np.maximum(bz, 0)
    iR = br.astype(np.int64)
    iz = bz.astype(np.int64)
    fR = br - iR
    fz = bz - iz
    hiR = iR >= nb - 1
    hiz = iz >= nb - 1
    iR = np.where(hiR, nb - 2, iR)
    fR = np.where(hiR, 1.0, fR)
    iz = np.where(hiz, nb - 2, iz)
    fz = np.where(hiz, 1.0, fz)
    return iR, iz, fR, fz


def disp_tables(p, cfg, t):
    """Target dispersion tables of type t on the force-grid nodes, as get_disp_rtp
    (disp_fields.c:66-198) interpolates them: FG_Disp_r at radii R_j for the spherical
    structures, FG_DispZ and FG_DispPhi at (R_j, z_k) for structure 2.
    Returns (beta parameter, dict of tables for GalicGPU.set_disp_tables)."""
    tovs = p.TypeOfVelocityStructure[t]
    beta = cfg.halo_beta if t == 1 else 0.0
    nodes = fg_nodes(p)
    if tovs >= 2:
        nb = p.FG_Nbin
        R = nodes[None, :].repeat(nb, 0).ravel()   # [k][j]: z row k, R column j
        z = nodes[:, None].repeat(nb, 1).ravel()
        d, _ = disp_at(p, t, tovs, beta, R, np.zeros_like(R), z, {})
        if tovs == 2:
            return beta, dict(disp_z=d[:, 1].copy(), disp_phi=d[:, 2].copy())
        return beta, dict(tilt_r=d[:, 0].copy(), tilt_theta=d[:, 1].copy(), disp_phi=d[:, 2].copy())
    d, _ = disp_at(p, t, tovs, beta, nodes, np.zeros_like(nodes), np.zeros_like(nodes), {})
    return beta, dict(disp_r=d[:, 0].copy())


def vstream_table(p, t, cfg):
    if t != 2 or p.TypeOfVelocityStructure[2] < 2:
        return np.zeros(p.FG_Nbin * p.FG_Nbin)
    nodes = fg_nodes(p)
    _, vstr = disk_moments(p, nodes)
    return np.ascontiguousarray(np.tile(vstr[None, :], (p.FG_Nbin, 1)).ravel())


def interp_table(p, table, x, y, z):
    """Bilinear lookup on an FG table (forcegrid_get_cell, grid.c:15-55)."""
    nb = p.FG_Nbin
    R = np.sqrt(x * x + y * y)
    br = np.maximum(np.log(R / p.FG_Rmin + 1.0) / np.log(p.FG_Fac), 0)
    bz = np.maximum(np.log(np.abs(z) / p.FG_Rmin + 1.0) / np.log(p.FG_Fac), 0)
    iR = br.astype(np.int64)
    iz = bz.astype(np.int64)
    fR = br - iR
    fz = bz - iz
    hiR = iR >= nb - 1
    hiz = iz >= nb - 1
    iR = np.where(hiR, nb - 2, iR)
    fR = np.where(hiR, 1.0, fR)
    iz = np.where(hiz, nb - 2, iz)
    fz = np.where(hiz, 1.0, fz)
    T = table
    return (T[iz * nb + iR] * (1 - fR) * (1 - fz) + T[(iz + 1) * nb + iR] * (1 - fR) * fz +
            T[iz * nb + iR + 1] * fR * (1 - fz) + T[(iz + 1) * nb + iR + 1] * fR * fz)


# --------------------------------------------------------------- particles --

def make_workload(cfg, n=None, target_samples=None, seed=None):
    """Build every input of the hot path for a configuration."""
    if n is not None or seed is not None:
        d = dict(cfg.__dict__)
        if n is not None:
            d["n"] = n
        if seed is not None:
            d["seed"] = seed
        cfg = Config(**d)
    p = make_params(cfg)
    pot, dpotdr, dpotdz = make_tables(p, cfg)
    vstream = {t: vstream_table(p, t, cfg) for t in (1, 2, 3)}
    nh, nd, nb = int(p.NType[1]), int(p.NType[2]), int(p.NType[3])
    ptype = np.concatenate([np.full(nh, 1), np.full(nd, 2), np.full(nb, 3)]).astype(np.int32)
    n = len(ptype)
    mass = np.zeros(n)
    for t in (1, 2, 3):
        if p.NType[t] > 0:
            mass[ptype == t] = p.MType[t] / p.NType[t]
    ids = np.arange(n, dtype=np.uint64)
    pos = np.zeros((n, 3))
    for t in (1, 2, 3):
        sel = np.nonzero(ptype == t)[0]
        if len(sel):
            pos[sel] = sample_type(p, t, cfg.seed, t, ids[sel])
    acc, phi = model_field(p, pos[:, 0], pos[:, 1], pos[:, 2])
    vesc = np.sqrt(-2.0 * phi)
    a = np.sqrt((acc * acc).sum(-1))
    r = np.sqrt((pos * pos).sum(-1))
    tint = cfg.torbit_fac * 2 * np.pi * r / np.sqrt(r * a)   # set_particles.c:87
    disp = np.zeros((n, 4))
    vstr = np.zeros(n)
    cache = {}
    for t in (1, 2, 3):
        sel = ptype == t
        if sel.any():
            beta = cfg.halo_beta if t == 1 else 0.0
            d, _ = disp_at(p, t, p.TypeOfVelocityStructure[t], beta, pos[sel, 0], pos[sel, 1], pos[sel, 2], cache)
            disp[sel] = d
            vstr[sel] = interp_table(p, vstream[t], pos[sel, 0], pos[sel, 1], pos[sel, 2])
    disp = np.maximum(disp, p.LowerDispLimit)   # set_particles.c:137-155
    vel = initial_velocities(p, cfg, pos, disp, vstr, vesc, ptype)
    if target_samples is None:
        target_samples = cfg.target_samples or max(200000, 10 * n)
    targets = {}
    for t in (1, 2, 3):
        if p.NType[t] > 0:
            mt, er, et, ep, eq = energy_targets(p, cfg, t, target_samples, cfg.seed)
            targets[t] = [density_target(p, t), mt, er, et, ep, eq]
    return Workload(cfg=cfg, params=p, pot=pot, dpotdr=dpotdr, dpotdz=dpotdz, vstream=vstream, targets=targets,
                    pos=np.ascontiguousarray(pos), vel=np.ascontiguousarray(vel), mass=mass, tint=tint,
                    ptype=ptype, vesc=vesc, disp=disp, vstr=vstr)


def local_basis(pos, cylindrical):
    """Unit vectors (R, T, Q) of optimize() (main.c:383-405)."""
    x, y, z = pos[:, 0], pos[:, 1], pos[:, 2]
    phi = np.arctan2(y, x)
    if cylindrical:
        R = np.stack([np.cos(phi), np.sin(phi), np.zeros_like(phi)], -1)
        T = np.stack([-np.sin(phi), np.cos(phi), np.zeros_like(phi)], -1)
        Q = np.stack([np.zeros_like(phi), np.zeros_like(phi), np.ones_like(phi)], -1)
        return R, T, Q
    theta = np.arccos(z / np.sqrt(x * x + y * y + z * z))
    R = np.stack([np.sin(theta) * np.cos(phi), np.sin(theta) * np.sin(phi), np.cos(theta)], -1)
    T = np.stack([-np.cos(theta) * np.cos(phi), -np.cos(theta) * np.sin(phi), np.sin(theta)], -1)
    Q = np.stack([-np.sin(phi), np.cos(phi), np.zeros_like(phi)], -1)
    return R, T, Q


def initial_velocities(p, cfg, pos, disp, vstr, vesc, ptype):
    """Gaussian velocities with the Jeans dispersions (set_particles.c:163-257)."""
    n = len(ptype)
    vel = np.zeros((n, 3))
    tovs = np.array([p.TypeOfVelocityStructure[t] for t in range(4)])[ptype]
    cyl = tovs == 2
    Rv, Tv, Qv = local_basis(pos, False)
    Rc, Tc, Qc = local_basis(pos, True)
    todo = np.arange(n)
    attempt = 0
    while len(todo):
        g = gaussians(cfg.seed, 7, todo.astype(np.uint64), attempt)
        s = np.sqrt(disp[todo])
        v1 = g[0] * s[:, 0]
        v2 = g[1] * s[:, 1]
        v3 = g[2] * s[:, 2] + vstr[todo]
        v = np.sqrt(v1 * v1 + v2 * v2 + v3 * v3)
        ok = v < p.MaxVelInUnitsVesc * vesc[todo]
        c = cyl[todo]
        # spherical: vr along e_r, vtheta along e_theta, vphi along e_phi; cylindrical: vR, vz, vphi
        sph = v1[:, None] * Rv[todo] + v3[:, None] * Qv[todo] + v2[:, None] * Tv[todo]
        cy = v1[:, None] * Rc[todo] + v3[:, None] * Tc[todo] + v2[:, None] * Qc[todo]
        vv = np.where(c[:, None], cy, sph)
        vel[todo[ok]] = vv[ok]
        todo = todo[~ok]
        attempt += 1
        if attempt > 1000:
            raise RuntimeError("velocity rejection loop did not finish")
    return vel


def trial_velocities(w, idx, seed, rep):
    """Fresh trial velocities of optimize() (main.c:383-435) for particles idx,
    drawn from the counter-based RNG keyed by (seed, rep)."""
    idx = np.asarray(idx)
    p = w.params
    pos = w.pos[idx]
    tovs = np.array([p.TypeOfVelocityStructure[t] for t in range(4)])[w.ptype[idx]]
    cyl = tovs == 2
    Rv, Tv, Qv = local_basis(pos, False)
    Rc, Tc, Qc = local_basis(pos, True)
    s = np.sqrt(w.disp[idx])
    out = np.zeros((len(idx), 3))
    todo = np.arange(len(idx))
    attempt = 0
    while len(todo):
        g = gaussians(seed, 1000 + rep, idx[todo].astype(np.uint64), attempt)
        vr = g[0] * s[todo, 0]
        vt = g[1] * s[todo, 1]
        vq = w.vstr[idx[todo]] + g[2] * s[todo, 2]
        sph = vr[:, None] * Rv[todo] + vt[:, None] * Tv[todo] + vq[:, None] * Qv[todo]
        cy = vr[:, None] * Rc[todo] + vt[:, None] * Tc[todo] + vq[:, None] * Qc[todo]
        v = np.where(cyl[todo][:, None], cy, sph)
        ok = np.sqrt((v * v).sum(-1)) < p.MaxVelInUnitsVesc * w.vesc[idx[todo]]
        out[todo[ok]] = v[ok]
        todo = todo[~ok]
        attempt += 1
        if attempt > 100000:
            raise RuntimeError("iter > 100000 (main.c:432-433)")
    return out


def permutation(w, seed):
    """The fixed random order in which optimize_some_particles walks the particles
    (set_particles.c:298-303)."""
    u = uniforms(seed, 9, np.arange(w.n, dtype=np.uint64))[0]
    return np.argsort(u, kind="stable").astype(np.int32)
