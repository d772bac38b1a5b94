"""Reader for muMatScale control files (read_ctrl.c:235-812 format).

Format: one ``Key value...`` per line, keys case-insensitive, lines whose
first token starts with ``#`` or ``%`` are comments, trailing text after the
expected values is ignored (sscanf reads only the expected fields), unknown
keys are ignored (read_ctrl.c:794-800), and ``GeoFileName`` /
``MatFileName`` / ``AlloyPropsFile0`` include another file (read_ctrl.c:623-631).
Defaults are those of set_defaults (read_ctrl.c:30-108).

The result is a :class:`Config` holding the step parameters (``mmca_params``)
and the nucleation set-up (``mmca_nucleation_cfg``) for one subblock.
"""
import ctypes as C
import math
import os
import re
import time

from . import abi

_INT = re.compile(r"^\s*[+-]?\d+")
_FLT = re.compile(r"^\s*[+-]?(?:\d+\.?\d*(?:[eE][+-]?\d+)?|\.\d+(?:[eE][+-]?\d+)?|inf|nan)", re.I)


def _scan(rest, fmt):
    """Minimal sscanf for %d/%u/%lu (``i``) and %lf (``f``) conversions."""
    out = []
    for f in fmt:
        m = (_INT if f == "i" else _FLT).match(rest)
        if not m:
            raise ValueError("cannot parse %r" % rest)
        tok = m.group(0)
        out.append(int(tok) if f == "i" else float(tok))
        rest = rest[m.end():]
    return out


class Config:
    """Settings of a control file with the reference's defaults."""

    def __init__(self):
        # set_defaults (read_ctrl.c:30-108)
        self.basefilename = "output"
        self.origin = [0.0, 0.0, 0.0]
        self.pad_bndy = [1, 1, 1]
        self.nsub = [1, 1, 1]
        self.sbdim = [0, 0, 0]
        self.cell_size = float(C.c_float(0.01).value)
        self.courant_limit = 0.2
        self.t_init = float(C.c_float(100.0).value)
        self.cool_rate = 0.0
        self.liquidus_temp = float(C.c_float(1180.0).value)
        self.solidus_temp = float(C.c_float(1120.0).value)
        self.random_seed = int(time.time()) & 0xFFFFFFFF
        self.fs_finish = 1.0
        self.finish_time = 0.0
        self.dt = float(C.c_float(0.1).value)
        self.screenpfreq = 1
        self.data_write_freq = 0
        self.data_write_start = 0
        self.data_format = "NONE"        # DataFormat: NONE, VTK or XDMF
        self.randomize_grains = 1
        self.bulk_density = 0.0
        self.bulk_mean = 0.0
        self.bulk_sigma = 0.0
        self.oriented = 0
        self.surf_density = 0.0
        self.surf_mean = 0.0
        self.surf_sigma = 0.0
        self.max_total_grains = 1000000
        self.do_surface = 1
        self.m_solute = [0.0, 0.0, -2.0]
        self.c_init = 5.0
        self.d_sol = 1e-13
        self.d_liq = 1e-9
        self.part_coef = 0.3
        self.fsgrow = 0.0
        self.thermal_ana = 0
        self.am_layers = 1
        self.am_tracks = 1
        self.gradient = 0.0
        self.grad_coef = 0.0
        self.velocity = 0.0
        self.velo_coef = 0.0
        self.laser = dict(power=0.0, speed=0.0, spot=0.0, absorp=0.0, hatch=0.0,
                          sub_init_temp=0.0, th_conductivity=0.0, density=0.0,
                          specific_heat=0.0, start=[0.0, 0.0, 0.0])
        self.fixed = []      # (x, y, z, ax, ay, az, threshold) in file order
        self.volumes = []    # (density/100, minx, miny, minz, maxx, maxy, maxz, threshold)

    # ---- parsing ---------------------------------------------------------
    def read(self, path):
        base = os.path.dirname(os.path.abspath(path))
        with open(path) as fh:
            for line in fh:
                line = line[:256]
                toks = line.split(None, 1)
                if not toks or toks[0][0] in "#%":
                    continue
                key = toks[0].lower()
                rest = toks[1] if len(toks) > 1 else ""
                self._apply(key, rest, base)
        return self

    def _apply(self, key, rest, base):
        f = lambda n=1: _scan(rest, "f" * n)
        i = lambda n=1: _scan(rest, "i" * n)
        if key == "basefilename":
            self.basefilename = rest.split()[0]
        elif key == "nsubblocks":
            self.nsub = i(3)
        elif key == "ncellspersb":
            self.sbdim = i(3)
        elif key == "cellsize":
            self.cell_size = f()[0]
        elif key == "courantlimit":
            self.courant_limit = f()[0]
        elif key == "facectrl":
            v = i(6)
            if v[0] != v[1] or v[2] != v[3] or v[4] != v[5]:
                raise ValueError("FaceCtrl: both faces of an axis must match (read_ctrl.c:293-296)")
            self.pad_bndy = [v[0], v[2], v[4]]
        elif key == "fixednuc":
            self.fixed.append(tuple(f(7)))
        elif key == "nucvolume":
            v = f(8)
            self.volumes.append((v[0] / 100.0,) + tuple(v[1:]))
        elif key == "maxgraindensity":
            self.bulk_density = f()[0]
        elif key == "maxgraindensitysurf":
            self.surf_density = f()[0]
        elif key == "maxtotgrains":
            self.max_total_grains = i()[0]
        elif key in ("mouldnuc", "moldnuc"):
            self.do_surface = i()[0]
        elif key in ("gngausscentre", "gngausscenter"):
            self.bulk_mean = f()[0]
        elif key == "gngausssigma":
            self.bulk_sigma = f()[0]
        elif key in ("gngausscentresurf", "gngausscentersurf"):
            self.surf_mean = f()[0]
        elif key == "gngausssigmasurf":
            self.surf_sigma = f()[0]
        elif key == "gnoriented":
            self.oriented = i()[0]
        elif key == "liquidustemp":
            self.liquidus_temp = f()[0]
        elif key == "solidustemp":
            self.solidus_temp = f()[0]
        elif key == "cinit":
            self.c_init = f()[0]
        elif key == "m_solute0":
            self.m_solute = f(3)
        elif key == "part_coef0":
            self.part_coef = f(3)[2]
        elif key == "dsol":
            self.d_sol = f()[0]
        elif key == "dliq":
            self.d_liq = f()[0]
        elif key == "initialtemperature":
            self.t_init = f()[0]
        elif key == "velo_coef":
            self.velo_coef = f()[0]
        elif key == "grad_coef":
            self.grad_coef = f()[0]
        elif key == "gradient":
            self.gradient = f()[0]
        elif key == "velocity":
            self.velocity = f()[0]
        elif key == "coolingrate":
            self.cool_rate = f()[0]
        elif key == "fsgrow":
            self.fsgrow = f()[0]
        elif key == "timestep":
            self.dt = f()[0]
        elif key == "finishtime":
            self.finish_time = f()[0]
        elif key == "fsfinish":
            self.fs_finish = f()[0]
        elif key == "screenpfreq":
            self.screenpfreq = i()[0]
        elif key == "datawritefreq":
            self.data_write_freq = i()[0]
        elif key == "datawritestart":
            self.data_write_start = i()[0]
        elif key == "dataformat":
            fmt = rest.split()[0].upper()
            if fmt not in ("VTK", "XDMF"):
                raise ValueError("Unsupported data file type: %s" % rest.split()[0])
            self.data_format = fmt
        elif key == "randomizegrains":
            self.randomize_grains = i()[0]
        elif key in ("geofilename", "matfilename", "alloypropsfile0"):
            inc = rest.split()[0]
            self.read(inc if os.path.isabs(inc) else os.path.join(base, inc))
        elif key == "randseedval":
            self.random_seed = i()[0] & 0xFFFFFFFF
        elif key == "analyticalthermal":
            self.thermal_ana = i()[0]
        elif key == "laserpower":
            self.laser["power"] = f()[0]
        elif key == "laserspeed":
            self.laser["speed"] = f()[0]
        elif key == "laserspotsize":
            self.laser["spot"] = f()[0]
        elif key == "laserabsorption":
            self.laser["absorp"] = f()[0]
        elif key == "laserhatchspacing":
            self.laser["hatch"] = f()[0]
        elif key == "subinitialtemp":
            self.laser["sub_init_temp"] = f()[0]
        elif key == "thermalconductivity":
            self.laser["th_conductivity"] = f()[0]
        elif key == "density":
            self.laser["density"] = f()[0]
        elif key == "specificheat":
            self.laser["specific_heat"] = f()[0]
        elif key == "laserstartlocation":
            self.laser["start"] = f(3)
        elif key == "nooflayers":
            self.am_layers = i()[0]
        elif key == "nooftracks":
            self.am_tracks = i()[0]
        # every other key is accepted and ignored, as in the reference

    # ---- derived ABI structures -------------------------------------------
    @property
    def gdim(self):
        return [self.nsub[a] * self.sbdim[a] for a in range(3)]

    def params(self, check_negative_ce=True, precision=abi.PREC_FP64, sb_coords=(0, 0, 0),
               whole_domain=True):
        """mmca_params of the whole domain as one block (whole_domain), or of
        the subblock at sb_coords."""
        p = abi.Params()
        dims = self.gdim if whole_domain else self.sbdim
        p.nx, p.ny, p.nz = dims
        p.pad_bndy[:] = self.pad_bndy
        p.sb_coords[:] = list(sb_coords)
        p.gdim[:] = self.gdim
        p.cell_size = self.cell_size
        p.origin[:] = self.origin
        p.dt = self.dt
        p.courant_limit = self.courant_limit
        p.c_init = self.c_init
        p.liquidus_temp = self.liquidus_temp
        p.solidus_temp = self.solidus_temp
        p.m_solute[:] = self.m_solute
        p.part_coef = self.part_coef
        p.d_sol = self.d_sol
        p.d_liq = self.d_liq
        p.fsgrow = self.fsgrow
        if self.thermal_ana:
            p.temp_model = abi.TEMP_AM_BEAM
        elif self.gradient != 0:
            p.temp_model = abi.TEMP_GRADIENT
        else:
            p.temp_model = abi.TEMP_UNIFORM
        p.am_layers = self.am_layers
        p.am_tracks = self.am_tracks
        p.gnsbx = self.nsub[0] if not whole_domain else 1
        if whole_domain and self.nsub[0] != 1:
            # the beam track length uses gnsbx * nx (temperature.c:68)
            p.gnsbx = 1
        p.t_init = self.t_init
        p.cool_rate = self.cool_rate
        p.gradient = self.gradient
        p.grad_coef = self.grad_coef
        p.velocity = self.velocity
        p.velo_coef = self.velo_coef
        L = self.laser
        p.laser_power = L["power"]
        p.laser_speed = L["speed"]
        p.laser_spot = L["spot"]
        p.laser_absorp = L["absorp"]
        p.laser_hatch = L["hatch"]
        p.sub_init_temp = L["sub_init_temp"]
        p.th_conductivity = L["th_conductivity"]
        p.density = L["density"]
        p.specific_heat = L["specific_heat"]
        p.laser_start[:] = L["start"]
        p.max_total_grains = self.max_total_grains
        p.check_negative_ce = 1 if check_negative_ce else 0
        p.precision = precision
        p.timestep = 0
        return p

    def nucleation(self, rank=0):
        """mmca_nucleation_cfg; the returned object keeps its arrays alive."""
        cfg = abi.NucleationCfg()
        cfg.random_seed = self.random_seed & 0xFFFFFFFF
        cfg.rank = rank
        cfg.oriented = self.oriented
        cfg.do_surface = self.do_surface
        cfg.bulk_density = self.bulk_density
        cfg.bulk_mean = self.bulk_mean
        cfg.bulk_sigma = self.bulk_sigma
        cfg.surf_density = self.surf_density
        cfg.surf_mean = self.surf_mean
        cfg.surf_sigma = self.surf_sigma
        fixed = (abi.FixedNuc * max(1, len(self.fixed)))(*[abi.FixedNuc(*t) for t in self.fixed])
        vols = (abi.NucVolume * max(1, len(self.volumes)))(*[abi.NucVolume(*t) for t in self.volumes])
        cfg.n_fixed = len(self.fixed)
        cfg.n_volumes = len(self.volumes)
        cfg.fixed = C.cast(fixed, C.POINTER(abi.FixedNuc))
        cfg.volumes = C.cast(vols, C.POINTER(abi.NucVolume))
        cfg._keep = (fixed, vols)
        return cfg

    def total_non_mold_volume(self):
        """totalNonMoldVolume_int (temperature.c:298-303)."""
        g = self.gdim
        return (g[0] * g[1] * g[2]) * math.pow(self.cell_size, 3)


def write_ctrl(cfg, path, nsub=None):
    """Write `cfg` as a control file the reference reads (read_ctrl.c keys),
    optionally with another subblock decomposition `nsub` of the same
    domain.  Round-trips through Config.read."""
    nsub = list(cfg.nsub) if nsub is None else list(nsub)
    g = cfg.gdim
    if any(g[a] % nsub[a] for a in range(3)):
        raise ValueError("domain %s does not split into %s subblocks" % (g, nsub))
    sb = [g[a] // nsub[a] for a in range(3)]
    r = repr
    L = cfg.laser
    lines = [
        "BaseFileName %s" % cfg.basefilename,
        "NSubblocks %d %d %d" % tuple(nsub),
        "NCellsPerSB %d %d %d" % tuple(sb),
        "FaceCtrl %d %d %d %d %d %d" % (cfg.pad_bndy[0], cfg.pad_bndy[0], cfg.pad_bndy[1], cfg.pad_bndy[1],
                                        cfg.pad_bndy[2], cfg.pad_bndy[2]),
        "CellSize %s" % r(cfg.cell_size),
        "CourantLimit %s" % r(cfg.courant_limit),
        "FSGrow %s" % r(cfg.fsgrow),
        "TimeStep %s" % r(cfg.dt),
        "RandSeedVal %d" % cfg.random_seed,
        "FinishTime %s" % r(cfg.finish_time),
        "FsFinish %s" % r(cfg.fs_finish),
        "ScreenPFreq %d" % cfg.screenpfreq,
        "DataWriteFreq %d" % cfg.data_write_freq,
        "RandomizeGrains %d" % cfg.randomize_grains,
        "LiquidusTemp %s" % r(cfg.liquidus_temp),
        "SolidusTemp %s" % r(cfg.solidus_temp),
        "Cinit %s" % r(cfg.c_init),
        "m_solute0 %s %s %s" % tuple(r(v) for v in cfg.m_solute),
        "part_coef0 0 0 %s" % r(cfg.part_coef),
        "Dsol %s" % r(cfg.d_sol),
        "Dliq %s" % r(cfg.d_liq),
        "InitialTemperature %s" % r(cfg.t_init),
        "CoolingRate %s" % r(cfg.cool_rate),
        "Gradient %s" % r(cfg.gradient),
        "Grad_Coef %s" % r(cfg.grad_coef),
        "Velocity %s" % r(cfg.velocity),
        "Velo_Coef %s" % r(cfg.velo_coef),
        "MaxGrainDensity %s" % r(cfg.bulk_density),
        "GNGaussCentre %s" % r(cfg.bulk_mean),
        "GNGaussSigma %s" % r(cfg.bulk_sigma),
        "MaxGrainDensitySurf %s" % r(cfg.surf_density),
        "GNGaussCentreSurf %s" % r(cfg.surf_mean),
        "GNGaussSigmaSurf %s" % r(cfg.surf_sigma),
        "MaxTotGrains %d" % cfg.max_total_grains,
        "MouldNuc %d" % cfg.do_surface,
        "GNOriented %d" % cfg.oriented,
        "AnalyticalThermal %d" % cfg.thermal_ana,
        "LaserPower %s" % r(L["power"]),
        "LaserSpeed %s" % r(L["speed"]),
        "LaserSpotSize %s" % r(L["spot"]),
        "LaserAbsorption %s" % r(L["absorp"]),
        "LaserHatchSpacing %s" % r(L["hatch"]),
        "SubInitialTemp %s" % r(L["sub_init_temp"]),
        "ThermalConductivity %s" % r(L["th_conductivity"]),
        "Density %s" % r(L["density"]),
        "SpecificHeat %s" % r(L["specific_heat"]),
        "LaserStartLocation %s %s %s" % tuple(r(v) for v in L["start"]),
        "NoOfLayers %d" % cfg.am_layers,
        "NoOfTracks %d" % cfg.am_tracks,
    ]
    if cfg.data_format != "NONE":
        lines.append("DataFormat %s" % cfg.data_format)
    if cfg.data_write_start:
        # the reference also takes this value as its checkpoint frequency
        lines.append("DataWriteStart %d" % cfg.data_write_start)
    for f in cfg.fixed:
        lines.append("FixedNuc %s %s %s %s %s %s %s" % tuple(r(v) for v in f))
    for v in cfg.volumes:
        lines.append("NucVolume %s %s %s %s %s %s %s %s" % ((r(v[0] * 100.0),) + tuple(r(x) for x in v[1:])))
    with open(path, "w") as fh:
        fh.write("\n".join(lines) + "\n")


def read_ctrl(path):
    return Config().read(path)
