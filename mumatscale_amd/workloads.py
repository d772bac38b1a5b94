"""Synthetic workloads of the CA step, as control-file settings (`Config`).

Alloy of every workload: c0 = 4 wt%, k = 0.14, liquidus slope -2.6 K/wt%,
T_liq = 933 K, D_l = 3e-9 m^2/s, D_s = 1e-12 m^2/s, dx = 1 um, FSGrow 0.66.
"""
import math

import numpy as np

from . import abi
from .ctrl import Config

DT_EXPLICIT = 0.9 * 1.0e-6 ** 2 / (6 * 3.0e-9)     # 0.9 dx^2 / (6 D_l) = 5e-5 s


def _alloy(c, nx, ny, nz):
    c.nsub = [1, 1, 1]
    c.sbdim = [nx, ny, nz]
    c.pad_bndy = [1, 1, 1]
    c.cell_size = 1.0e-6
    c.fsgrow = 0.66
    c.dt = DT_EXPLICIT
    c.liquidus_temp = 933.0
    c.solidus_temp = 850.0
    c.c_init = 4.0
    c.m_solute = [0.0, 0.0, -2.6]
    c.part_coef = 0.14
    c.d_sol = 1.0e-12
    c.d_liq = 3.0e-9
    return c


def directional_config(nx=256, ny=256, nz=512):
    """Directional solidification: frozen T = T_liq - 5 K + G (z - V t),
    G = 1e4 K/m, V = 1e-3 m/s; surface nucleation 1e10 m^-2 (mean 1 K, sigma
    0.2 K) on every padded face, bulk 1e13 m^-3 (mean 5 K, sigma 0.5 K),
    random orientations from seed 1."""
    c = _alloy(Config(), nx, ny, nz)
    c.random_seed = 1
    c.t_init = 933.0 - 5.0
    c.gradient = 1.0e4
    c.velocity = 1.0e-3
    c.bulk_density = 1.0e13
    c.bulk_mean = 5.0
    c.bulk_sigma = 0.5
    c.surf_density = 1.0e10
    c.surf_mean = 1.0
    c.surf_sigma = 0.2
    c.oriented = 1
    c.max_total_grains = 1000000
    return c


def ensemble_configs(n=96, seed=3):
    """96 independent domains: sides drawn from {32,48,64,96,128} per axis,
    cooling rate log-uniform in [1, 1000] K/s, initial undercooling uniform
    in [1, 5] K, bulk nucleation 1e14 m^-3 (mean 3 K, sigma 0.5 K) with a
    distinct seed per domain; alloy and dt of the single-grain physics."""
    rng = np.random.default_rng(seed)
    sides = rng.choice([32, 48, 64, 96, 128], size=(n, 3))
    rates = 10.0 ** rng.uniform(0.0, 3.0, size=n)
    under = rng.uniform(1.0, 5.0, size=n)
    out = []
    for i in range(n):
        c = directional_config(*(int(v) for v in sides[i]))
        c.gradient = 0.0
        c.velocity = 0.0
        c.cool_rate = float(rates[i])
        c.t_init = 933.0 - float(under[i])
        c.surf_density = 0.0
        c.bulk_density = 1.0e14
        c.bulk_mean = 3.0
        c.bulk_sigma = 0.5
        c.random_seed = 1000 + i
        c.max_total_grains = 100000
        out.append(c)
    return out


def _zxz_degrees(R):
    """Proper Euler (z-x-z) angles in degrees of the rotation R, in the
    argument order of grain.c:314-336.  That code builds the z-x-z matrix
    except for entry [0][0] (grain.c:326 has s2 s1 s0 where the rotation has
    s2 c1 s0), so the matrix the CA uses from these angles is R up to that
    entry; the build keeps the reference's formula."""
    a1 = math.acos(max(-1.0, min(1.0, R[2][2])))
    a0 = math.atan2(R[0][2], -R[1][2])
    a2 = math.atan2(R[2][0], R[2][1])
    return tuple(math.degrees(a) for a in (a0, a1, a2))


def _quat_matrix(q):
    w, x, y, z = q
    return [[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
            [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
            [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]]


def seeded_grains_config(n=128, nseeds=16, seed=4, dt_scale=0.5, ny=None, nz=None):
    """Equiaxed grains of the single-grain physics: uniform T starting 2 K
    below the liquidus and cooling at 10 K/s, zero-flux walls, no random
    nucleation, `nseeds` FixedNuc seeds at distinct random cells (at least
    2 cells from the walls) with the Euler angles of uniform-random
    rotations (normalised Gaussian quaternions), all from numpy seed `seed`.  dt = dt_scale x the
    explicit limit: the literal limit makes ce negative within ~50 steps
    (an unstable partitioning update), half of it runs stably."""
    ny = n if ny is None else ny
    nz = n if nz is None else nz
    c = _alloy(Config(), n, ny, nz)
    c.dt = DT_EXPLICIT * dt_scale
    c.random_seed = seed
    c.t_init = 933.0 - 2.0
    c.cool_rate = 10.0
    c.bulk_density = 0.0
    c.surf_density = 0.0
    c.oriented = 1
    c.max_total_grains = 1000
    rng = np.random.default_rng(seed)
    dims = (n - 4, ny - 4, nz - 4)
    cells = rng.choice(dims[0] * dims[1] * dims[2], size=nseeds, replace=False)
    quats = rng.normal(size=(nseeds, 4))
    for cell, q in zip(cells, quats):
        q = q / np.linalg.norm(q)
        i = int(cell % dims[0]) + 2
        j = int((cell // dims[0]) % dims[1]) + 2
        k = int(cell // (dims[0] * dims[1])) + 2
        ax, ay, az = _zxz_degrees(_quat_matrix(q))
        c.fixed.append(((i + 0.5) * c.cell_size, (j + 0.5) * c.cell_size, (k + 0.5) * c.cell_size,
                        ax, ay, az, 0.0))
    return c


def am_track_config(nx=1024, ny=512, nz=256, dt_scale=0.5):
    """Additive-manufacturing single track over a polycrystalline substrate:
    dx = 0.5 um, the reference's Gaussian-beam temperature model
    (temperature.c:62-123) with power 200 W, absorptivity 0.35, speed
    0.8 m/s, spot 130 um, conductivity 20 W/m K, density x specific heat
    4e6 J/m^3 K (diffusivity 5e-6 m^2/s), substrate at 300 K; the beam starts
    at the -x face, mid-width, on the top surface.  Bulk nucleation as in
    the directional workload (1e13 m^-3, mean 5 K, sigma 0.5 K, seed 2), no
    surface nucleation.  dt = dt_scale x 0.9 dx^2 / (6 D_l)."""
    c = _alloy(Config(), nx, ny, nz)
    c.cell_size = 0.5e-6
    c.dt = 0.9 * c.cell_size ** 2 / (6 * c.d_liq) * dt_scale
    c.random_seed = 2
    c.thermal_ana = 1
    c.laser.update(power=200.0, speed=0.8, spot=130e-6, absorp=0.35, hatch=50e-6, sub_init_temp=300.0,
                   th_conductivity=20.0, density=8000.0, specific_heat=500.0,
                   start=[0.0, ny * c.cell_size / 2, nz * c.cell_size])
    c.am_layers = 1
    c.am_tracks = 1
    c.t_init = 300.0
    c.bulk_density = 1.0e13
    c.bulk_mean = 5.0
    c.bulk_sigma = 0.5
    c.surf_density = 0.0
    c.oriented = 1
    c.max_total_grains = 1000000
    return c


def voronoi_substrate(cfg, ngrains=2000, seed=2):
    """A substrate filling the block: the interior cells of each Voronoi
    region of `ngrains` uniform-random seed points form one grain with a
    uniform-random orientation (normalised Gaussian quaternion), all from
    numpy seed `seed`.  Returns (gr, records): the padded int32 gr field and
    the grain records 0..ngrains (record 0 is the liquid).  The nearest seed
    is found on the GPU with torch."""
    import torch
    nx, ny, nz = (cfg.nsub[a] * cfg.sbdim[a] for a in range(3))
    rng = np.random.default_rng(seed)
    pts = rng.uniform(size=(ngrains, 3)) * np.array([nx, ny, nz])
    quats = rng.normal(size=(ngrains, 4))
    dev = torch.device("cuda")
    P = torch.from_numpy(pts).to(dev)
    pn = (P * P).sum(1)
    gr = np.zeros((nz + 2, ny + 2, nx + 2), np.int32)
    xs = torch.arange(nx, device=dev, dtype=torch.float64) + 0.5
    ys = torch.arange(ny, device=dev, dtype=torch.float64) + 0.5
    Y, X = torch.meshgrid(ys, xs, indexing="ij")
    rows = max(1, (1 << 15) // nx)                   # 32k cells x ngrains distances per chunk
    for k in range(nz):
        plane = torch.empty((ny, nx), dtype=torch.int32, device=dev)
        for j0 in range(0, ny, rows):
            xq, yq = X[j0:j0 + rows].reshape(-1), Y[j0:j0 + rows].reshape(-1)
            Q = torch.stack([xq, yq, torch.full_like(xq, k + 0.5)], 1)
            d = pn[None, :] - 2.0 * (Q @ P.T)        # |q|^2 is the same for every seed
            plane[j0:j0 + rows] = (d.argmin(1) + 1).to(torch.int32).reshape(-1, nx)
        gr[k + 1, 1:-1, 1:-1] = plane.cpu().numpy()
    recs = [abi.Grain()]
    for i in range(ngrains):
        q = quats[i] / np.linalg.norm(quats[i])
        R = _quat_matrix(q)
        g = abi.Grain()
        g.nuc_x, g.nuc_y, g.nuc_z = (float(v) * cfg.cell_size for v in pts[i])
        g.nuc_timestep = 0
        g.nuc_threshold = 0.0
        for a in range(3):
            for b in range(3):
                g.rotmat[a][b] = R[a][b]
        g.rotang[0], g.rotang[1], g.rotang[2] = (math.radians(v) for v in _zxz_degrees(R))
        recs.append(g)
    return gr, recs


def apply_substrate(sb, gr, records):
    """Make every interior cell of `sb` solid (fs = 1) in the grain of `gr`."""
    sb.set_grains(records)
    sb.upload(abi.FIELD_GR, gr)
    fs = np.zeros(gr.shape)
    fs[1:-1, 1:-1, 1:-1] = 1.0
    sb.upload(abi.FIELD_FS, fs)
