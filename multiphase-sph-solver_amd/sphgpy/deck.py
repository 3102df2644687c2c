"""Input decks of the multi-phase SPH solver: parameters, materials and particles.

A ``Deck`` holds what the reference reads from ``input.dat`` (grammar: CPP_2-initialf.cpp:22-1018,
sections FL PR XP AT GF IP BP WP AP MS <model blocks> SM PA) and can write that file, so the
same seeded deck drives the reference solver and this library. The generators build the
workloads the benchmarks are quoted on (2D dry granular column, 3D soil-water column, 3D slope
with pile, 2D random-field slope), each scalable to test sizes.
"""
from __future__ import annotations

import copy
import math
from dataclasses import dataclass, field
from typing import Dict, List, Optional

import numpy as np

SOIL_FIELDS = ("e poi eps0 dens damp1 damp2 cop ramda kapa zmf ann bnn beta sitas sitar epsr epss "
               "sd sw kse c1 c2 c3 fai c ocr sme0 mzm azm r0 bl bz zet vpu porosity fs etao g0 ds "
               "damp_coe").split()
FLUID_FIELDS = "vis vpu dens paa0 porosity fai c bhtype press_type acc_type pre_coe strain_min".split()
STRUCT_FIELDS = "dens eps0 e poi vpu porosity cop".split()

# model index -> input.dat block tag and the per-layer field lines (CPP_2-initialf.cpp:487-795)
MODEL_BLOCKS = {
    0: ("EM", [["kapa", "poi", "dens"], ["damp1", "damp2", "cop", "eps0"], ["e", "ds"]]),
    1: ("SB", [["poi", "dens", "eps0", "zmf", "sme0", "fs"], ["ramda", "kapa", "damp1", "damp2", "cop"],
               ["ann", "ocr"], ["e", "ds"]]),
    2: ("DP", [["kapa", "poi", "dens", "c", "fai", "fs"], ["damp1", "damp2", "cop", "eps0"], ["e", "ds"]]),
    3: ("UM", [["poi", "dens", "eps0", "zmf", "sme0", "fs"], ["ramda", "kapa", "damp1", "damp2", "cop"],
               ["ann", "bnn", "beta", "epss", "epsr", "sitas", "sitar"],
               ["sd", "sw", "kse", "c1", "c2", "c3"], ["e", "ds"]]),
    4: ("AM", [["poi", "dens", "eps0", "zmf", "sme0", "fs"], ["ramda", "kapa", "damp1", "damp2", "cop"],
               ["mzm", "azm", "r0", "bl", "bz", "zet"], ["e", "ds"]]),
    5: ("BU", [["kapa", "poi", "dens", "c", "fai", "beta", "fs"], ["damp1", "damp2", "cop", "eps0"], ["e", "ds"]]),
    6: ("PP", [["kapa", "poi", "dens"], ["damp1", "damp2", "cop", "eps0"], ["e", "ds"]]),
    7: ("VE", [["kapa", "poi", "dens", "c"], ["damp1", "damp2", "cop", "eps0"], ["e", "ds"]]),
    8: ("FC", [["dens", "c", "eps0", "cop", "sme0"], ["g0", "poi", "zmf", "ramda", "etao"],
               ["zet", "ann", "bnn", "damp1", "damp2"], ["ds"]]),
    9: ("EP", [["kapa", "poi", "dens", "c", "fai", "fs"], ["damp1", "damp2", "cop", "eps0"], ["e", "ds"]]),
}
MODEL_FLAG_NAMES = ["flagem", "flagsb", "flagdp", "flagum", "flagam", "flagbui", "flagpl", "flagve",
                    "flagfocc", "flagep"]


def _g(v) -> str:
    """Shortest round-tripping decimal of a double (the reference parses it back with %lf)."""
    if isinstance(v, (int, np.integer)):
        return str(int(v))
    return repr(float(v))


def default_soil() -> Dict[str, float]:
    """Para_Soil::Para_Soil (Class_ParticleProperty.cpp:30-83)."""
    d = {k: 0.0 for k in SOIL_FIELDS}
    d["fs"] = 1.0
    d["ds"] = 1.0
    return d


def default_fluid() -> Dict[str, float]:
    """Para_Fluid::Para_Fluid (Class_ParticleProperty.cpp:89-106)."""
    d = {k: 0.0 for k in FLUID_FIELDS}
    d["strain_min"] = 1.0
    for k in ("bhtype", "press_type", "acc_type"):
        d[k] = 0
    return d


@dataclass
class Deck:
    ndim: int = 2
    # Para_FL
    fla: int = 0
    fls: int = 1
    flr: int = 0
    fli: int = 0
    flv: int = 0
    fleob: int = 0
    flts: int = 0
    flspv: int = 0
    flknl: int = 1
    ntd: int = 1
    flbndy: int = 2
    # Para_Pro
    threshold_min: float = 0.1
    threshold_max: float = 0.12
    fop: int = 1
    inip: int = 0
    loop: int = 1
    step_regu: int = 1000000
    dttype: int = 1
    dr: float = 0.002
    dt: float = 1e-6
    xsph: tuple = (0.5, 0.5, 0.5)
    art_vis: tuple = ((0.0, 0.0), (0.1, 0.0), (0.0, 0.0))
    art_stre_coe: float = -1.0
    gravity: tuple = (0.0, -9.81, 0.0)
    # Para_Inter / Para_SSInter
    interc: float = 1.0
    permtype: int = 0
    pers: float = 0.0
    perl: float = 0.0
    perm: float = 0.0
    ss_type: int = 0
    ss_myu: float = 0.0
    ss_zeta: float = 0.5
    ss_acc_type: int = 1
    # Para_Boundary
    if_cal: int = 0
    if_move: int = 0
    coe_bndy: float = 0.1
    move_vx: tuple = (0.0, 0.0, 0.0)
    # materials
    water: List[Dict[str, float]] = field(default_factory=lambda: [default_fluid() for _ in range(3)])
    air: Dict[str, float] = field(default_factory=default_fluid)
    model_flags: List[int] = field(default_factory=lambda: [0] * 10)
    alphats: float = 0.0
    damp_coe: float = 0.0
    soil_layers: Dict[int, List[Dict[str, float]]] = field(default_factory=dict)
    structure: Dict[str, float] = field(default_factory=lambda: {k: 0.0 for k in STRUCT_FIELDS})
    # particles
    xp: np.ndarray = None            # (n, 3) float64
    vxp: np.ndarray = None           # (n, 3) float64
    type: np.ndarray = None          # (n,) int32
    matype: np.ndarray = None
    permtype_p: np.ndarray = None
    etype: np.ndarray = None
    name: str = "deck"
    # Para_Vibra (flv 1, 2): one period of vib_ttime in vib_nsteps loads, applied between
    # vib_sttime and vib_edtime; vib_loads (vib_nsteps, 3) go to vibra.dat
    vib_ttime: float = 0.0
    vib_nsteps: int = 0
    vib_sttime: float = 0.0
    vib_edtime: float = 0.0
    vib_loads: np.ndarray = None
    # Para_Rain (flr 1): (rtype, nrain, nhor, press_factor, drop_velocity)
    rain: tuple = None
    # Para_EC (fleob 1): stages (first, last, step); input.dat carries first and last only
    ecp: list = None

    @property
    def n(self) -> int:
        return int(self.xp.shape[0])

    def copy(self) -> "Deck":
        return copy.deepcopy(self)

    # ------------------------------------------------------------------ input.dat writer
    def write_input_dat(self, path: str) -> None:
        """Write the reference deck file (grammar of CPP_2-initialf.cpp:22-1018).

        Every line has at least three characters: the section scanners use "%c%c %*[^\\n]%*c",
        which swallows the following line after a one-character line."""
        L: List[str] = []

        def sec(tag, rows):
            L.append(f"{tag} section")
            L.append("A data")
            for r in rows:
                L.append(" ".join(_g(v) for v in r))
            L.append("E end")

        sec("FL", [[self.fla, self.fls, self.flr, self.fli, self.flv, self.fleob, self.flts, self.flspv,
                    self.flknl], [self.ntd, self.flbndy]])
        sec("PR", [[self.ndim, self.threshold_min, self.threshold_max],
                   [self.fop, self.inip, self.loop, self.n, self.step_regu],
                   [self.dttype, self.dr, self.dt]])
        sec("XP", [list(self.xsph)])
        sec("AT", [list(self.art_vis[0]), [self.art_vis[1][0], self.art_vis[1][1], self.art_stre_coe],
                   list(self.art_vis[2])])
        sec("GF", [list(self.gravity)])
        sec("IP", [[self.interc], [self.permtype, self.pers, self.perl, self.perm],
                   [self.ss_type, self.ss_myu, self.ss_zeta, self.ss_acc_type]])
        brow = [self.if_cal, self.if_move, self.coe_bndy]
        rows = [brow] + ([list(self.move_vx)] if self.if_move >= 1 else [])
        sec("BP", rows)
        w0 = self.water[0]
        sec("WP", [[w0["dens"], w0["vis"], w0["fai"], w0["c"], w0["vpu"], w0["press_type"], w0["acc_type"],
                    w0["bhtype"], w0["pre_coe"], w0["strain_min"]]] +
            [[w["dens"], w["vis"], w["fai"], w["c"], w["vpu"]] for w in self.water[1:]])
        if self.fla == 1:
            a = self.air
            sec("AP", [[a["dens"], a["paa0"], a["vis"], a["vpu"], a["press_type"], a["acc_type"], a["pre_coe"]]])
        if self.fls == 1:
            ms = list(self.model_flags)
            rows = [ms] + ([[self.alphats, self.damp_coe]] if self.flts >= 1 else [])
            sec("MS", rows)
            for model, layers in sorted(self.soil_layers.items()):
                tag, lines = MODEL_BLOCKS[model]
                rows = []
                for lay in layers:
                    for ln in lines:
                        rows.append([lay[k] for k in ln])
                sec(tag, rows)
        s = self.structure
        sec("SM", [[s["dens"], s["eps0"], s["cop"]], [s["e"], s["poi"]]])
        rows = []
        for i in range(self.n):
            if self.ndim == 2:
                rows.append([self.xp[i, 0], self.xp[i, 1], self.vxp[i, 0], self.vxp[i, 1],
                             int(self.type[i]), int(self.matype[i]), int(self.permtype_p[i]), int(self.etype[i])])
            else:
                rows.append([self.xp[i, 0], self.xp[i, 1], self.xp[i, 2], self.vxp[i, 0], self.vxp[i, 1],
                             self.vxp[i, 2], int(self.type[i]), int(self.matype[i]), int(self.permtype_p[i]),
                             int(self.etype[i])])
        sec("PA", rows)
        if self.flr == 1:
            sec("RP", [list(self.rain)])  # initial_pro, CPP_2-initialf.cpp:180-201
        if self.fleob == 1:
            sec("EC", [[len(self.ecp)]] + [[int(a), int(b)] for a, b, _ in self.ecp])  # :125-157
        if self.flv >= 1:
            # VL section (initial_pro, CPP_2-initialf.cpp:250-265) and the loads of vibra.dat
            # beside input.dat (initial_vib, :293-322)
            sec("VL", [[self.vib_ttime, self.vib_nsteps], [self.vib_sttime, self.vib_edtime]])
            import os
            with open(os.path.join(os.path.dirname(os.path.abspath(path)), "vibra.dat"), "w") as f:
                for row in np.asarray(self.vib_loads, dtype=np.float64).reshape(-1, 3):
                    f.write(" ".join(_g(float(v)) for v in row) + "\n")
        with open(path, "w") as f:
            f.write("\n".join(L) + "\n")

    # ------------------------------------------------------------------ material tables
    def soil_table(self) -> List[Dict[str, float]]:
        """pSoil[30] after initial_mat (CPP_2-initialf.cpp:487-852): block values, vpu=600
        (:797-802), damp_coe, porosity = eps0/(1+eps0) when fls==1, FOCC e=5e6 (:738)."""
        tab = [default_soil() for _ in range(30)]
        if self.fls == 1:
            for model, layers in self.soil_layers.items():
                for lay_i, lay in enumerate(layers[: self.model_flags[model]]):
                    d = tab[lay_i * 10 + model]
                    for ln in MODEL_BLOCKS[model][1]:
                        for k in ln:
                            d[k] = float(lay[k])
                    if model == 8:
                        d["e"] = 5.0e6
        # damp_coe is read from the MS section's second line, which exists only when flts >= 1
        damp = float(self.damp_coe) if (self.fls == 1 and self.flts >= 1) else 0.0
        for d in tab:
            d["vpu"] = 600.0
            d["damp_coe"] = damp
        if self.fls == 1:
            for d in tab:
                d["porosity"] = d["eps0"] / (1 + d["eps0"])
        return tab

    def fluid_tables(self):
        """pWater[3] and pAir after initial_mat (CPP_2-initialf.cpp:410-440, :803-835)."""
        water = [dict(w) for w in self.water]
        for k in ("press_type", "bhtype", "pre_coe", "strain_min"):
            water[1][k] = water[0][k]
            water[2][k] = water[0][k]
        air = dict(self.air) if self.fla == 1 else default_fluid()
        soil = self.soil_table()
        if self.fls == 0:
            por = 1.0
        else:
            por = 0.0
            for d in soil:
                por = max(por, d["porosity"])
        air["porosity"] = por
        for w in water:
            w["porosity"] = por
        for w in water:
            if w["dens"] <= 0.0:
                w["dens"] = 1000.0
        if air["dens"] <= 0.0:
            air["dens"] = 1.293
        return water, air

    def structure_table(self) -> Dict[str, float]:
        s = dict(self.structure)
        s["vpu"] = 600.0
        s["porosity"] = s["eps0"] / (1 + s["eps0"])
        return s

    def domain(self):
        """Cell_Con box and per-phase extents as input_par computes them (CPP_2-initialf.cpp:937-1017),
        including the fmax(water_min, x) slip of the phase maxima."""
        xp = self.xp
        bmax = np.full(3, -10000.0)
        bmin = np.full(3, 10000.0)
        for j in range(3):
            bmax[j] = max(bmax[j], float(xp[:, j].max()))
            bmin[j] = min(bmin[j], float(xp[:, j].min()))
        for j in range(self.ndim):
            bmax[j] = bmax[j] + 2.0 * self.dr
            bmin[j] = bmin[j] - 2.0 * self.dr
        ph = {}
        for t, nm in ((1, "water"), (2, "soil")):
            # per particle in order: mn = min(mn, x), then mx = max(mn, x) = x, so the maximum
            # slot ends up holding the last particle's coordinates
            mn = np.full(3, 100000.0)
            mx = np.full(3, -100000.0)
            idx = np.nonzero(self.type == t)[0]
            if idx.size:
                mn = np.minimum(mn, xp[idx].min(axis=0))
                mx = xp[idx[-1]].astype(np.float64).copy()
            ph[nm] = (mn, mx)
        return bmin, bmax, ph


# ---------------------------------------------------------------------- workload generators

def _jitter(rng, n, ndim, dx):
    j = np.zeros((n, 3))
    j[:, :ndim] = rng.uniform(-1e-6 * dx, 1e-6 * dx, size=(n, ndim))
    return j


def _lattice2d(ix0, ix1, iy0, iy1, dx):
    xs = (np.arange(ix0, ix1) + 0.5) * dx
    ys = (np.arange(iy0, iy1) + 0.5) * dx
    X, Y = np.meshgrid(xs, ys, indexing="ij")
    p = np.zeros((X.size, 3))
    p[:, 0] = X.ravel()
    p[:, 1] = Y.ravel()
    return p


def _lattice3d(ix0, ix1, iy0, iy1, iz0, iz1, dx):
    xs = (np.arange(ix0, ix1) + 0.5) * dx
    ys = (np.arange(iy0, iy1) + 0.5) * dx
    zs = (np.arange(iz0, iz1) + 0.5) * dx
    X, Y, Z = np.meshgrid(xs, ys, zs, indexing="ij")
    return np.stack([X.ravel(), Y.ravel(), Z.ravel()], axis=1)


def _assemble(deck: Deck, parts, rng):
    xp = np.concatenate([p[0] for p in parts])
    n = xp.shape[0]
    deck.xp = xp + _jitter(rng, n, deck.ndim, deck.dr)
    deck.vxp = np.zeros((n, 3))
    deck.type = np.concatenate([np.full(p[0].shape[0], p[1], np.int32) for p in parts])
    deck.matype = np.concatenate([np.full(p[0].shape[0], p[2], np.int32) for p in parts])
    deck.permtype_p = np.concatenate([np.full(p[0].shape[0], p[3] if len(p) > 3 else 0, np.int32)
                                      for p in parts])
    deck.etype = np.ones(n, np.int32)
    return deck


def dp_layer(c, fai, e, poi, dens, eps0=0.0, kapa=0.01, fs=1.0, cop=0.0):
    d = default_soil()
    d.update(kapa=kapa, poi=poi, dens=dens, c=c, fai=fai, fs=fs, damp1=0.0, damp2=0.0, cop=cop, eps0=eps0,
             e=e, ds=1.0)
    return d


def column_collapse_2d(nx=60, ny=120, dx=0.002, wall_h=130, floor_len=180, steps=5000, seed=1,
                       c_sound=600.0, xsph=0.5) -> Deck:
    """Plain FP64 2D dry granular column collapse: nx*ny soil lattice at (i+1/2)dx, Drucker-Prager
    c=0, phi=30, psi=0, E=1.8 MPa, nu=0.3, rho=2040; 3-layer rigid floor (x from -3dx to
    floor_len*dx) and left wall (y up to wall_h*dx); AV alpha=0.1; dt = 0.2 dx / c_sound;
    seeded +-1e-6 dx jitter. Defaults give 7,200 soil + 939 wall = 8,139 particles."""
    rng = np.random.default_rng(seed)
    d = Deck(ndim=2, dr=dx, dt=0.2 * dx / c_sound, loop=steps, fop=steps, gravity=(0.0, -9.81, 0.0),
             xsph=(xsph, xsph, xsph), art_vis=((0.0, 0.0), (0.1, 0.0), (0.0, 0.0)), flbndy=2,
             name=f"column2d_{nx}x{ny}")
    d.model_flags[2] = 1
    d.soil_layers[2] = [dp_layer(c=0.0, fai=30.0, e=1.8e6, poi=0.3, dens=2040.0)]
    soil = _lattice2d(0, nx, 0, ny, dx)
    floor = _lattice2d(-3, floor_len, -3, 0, dx)
    wall = _lattice2d(-3, 0, 0, wall_h, dx)
    return _assemble(d, [(soil, 2, 2), (floor, 0, 0), (wall, 0, 0)], rng)


def _box_walls_3d(nx, ny, nz_wall, dx, layers=3):
    """Floor under [0,nx)x[0,ny) plus four side walls up to nz_wall, `layers` thick."""
    floor = _lattice3d(-layers, nx + layers, -layers, ny + layers, -layers, 0, dx)
    xm = _lattice3d(-layers, 0, -layers, ny + layers, 0, nz_wall, dx)
    xp_ = _lattice3d(nx, nx + layers, -layers, ny + layers, 0, nz_wall, dx)
    ym = _lattice3d(0, nx, -layers, 0, 0, nz_wall, dx)
    yp = _lattice3d(0, nx, ny, ny + layers, 0, nz_wall, dx)
    return np.concatenate([floor, xm, xp_, ym, yp])


def soil_water_column_3d(nx=64, ny=32, nz_soil=64, nz_water=16, dx=0.005, steps=2000, seed=2,
                         c_water=30.0, eps0=0.6, perm=1e-4) -> Deck:
    """3D saturated soil-water column: nx*ny*nz_soil Drucker-Prager soil (c=1 kPa, phi=25) under an
    nx*ny*nz_water water layer; walls on four sides and a floor; Darcy drag (fli=1, permtype=1,
    pers=perm); water EOS with vpu=c_water. dt = 0.2 dx / 600."""
    rng = np.random.default_rng(seed)
    d = Deck(ndim=3, dr=dx, dt=0.2 * dx / 600.0, loop=steps, fop=steps, gravity=(0.0, 0.0, -9.81),
             fli=1, permtype=1, pers=perm, perl=perm, perm=perm, interc=1.0, flbndy=2,
             art_vis=((0.1, 0.0), (0.1, 0.0), (0.0, 0.0)), xsph=(0.5, 0.5, 0.5),
             name=f"soilwater3d_{nx}x{ny}x{nz_soil}+{nz_water}")
    d.model_flags[2] = 1
    d.soil_layers[2] = [dp_layer(c=1000.0, fai=25.0, e=1.8e6, poi=0.3, dens=2650.0, eps0=eps0, cop=perm)]
    w = default_fluid()
    por = eps0 / (1 + eps0)
    # Tait stiffness B = rho0 c^2 / 7 carried by the initial pressure prep (CPP_7-1-eos_CR.cpp:28-29)
    B = 1000.0 * c_water ** 2 / 7.0
    w.update(dens=1000.0, vis=1e-3, vpu=c_water, press_type=0, acc_type=0, bhtype=0,
             pre_coe=B / (por * 1000.0 * 9.81 * dx), strain_min=1.0)
    d.water = [dict(w), dict(w), dict(w)]
    soil = _lattice3d(0, nx, 0, ny, 0, nz_soil, dx)
    water = _lattice3d(0, nx, 0, ny, nz_soil, nz_soil + nz_water, dx)
    walls = _box_walls_3d(nx, ny, nz_soil + nz_water + 4, dx)
    return _assemble(d, [(soil, 2, 2), (water, 1, 0), (walls, 0, 0)], rng)


def _layer(model: int, **kw) -> Dict[str, float]:
    d = default_soil()
    base = dict(poi=0.3, dens=2000.0, eps0=0.6, kapa=0.04, fs=1.0, cop=1e-4, e=2.0e6, ds=1.0, sitas=1.0, sitar=0.2,
                sd=1.0, sw=1.0, c1=1.0, c2=1.0)
    base.update(kw)
    d.update(base)
    return d


def model_zoo_2d(nx=12, nbands=7, band_h=3, water_h=3, dx=0.002, steps=12, seed=5, art_stress=0.3,
                 step_regu=3, inip=3) -> Deck:
    """2D column of horizontal soil bands, one per constitutive model (elastic, SB Cam-Clay, DP,
    UMCC, AMCC, Bui DP, FOCC), under a water layer, with artificial stress, periodic
    strain/stress regularisation and initial geostress damping switched on: a deck that drives
    every per-particle branch of the step loop the benchmark decks leave idle."""
    rng = np.random.default_rng(seed)
    d = Deck(ndim=2, dr=dx, dt=0.2 * dx / 600.0, loop=steps, fop=1, gravity=(0.0, -9.81, 0.0),
             art_vis=((0.1, 0.0), (0.1, 0.0), (0.0, 0.0)), art_stre_coe=art_stress, step_regu=step_regu,
             inip=inip, flbndy=2, fli=1, permtype=1, pers=1e-4, perl=1e-3, perm=5e-4, interc=1.0, damp_coe=0.05,
             name="model_zoo_2d")
    models = [0, 1, 2, 3, 4, 5, 8][:nbands]
    d.model_flags = [1 if m in models else 0 for m in range(10)]
    d.soil_layers = {
        0: [_layer(0, e=3.0e6)],
        1: [_layer(1, zmf=1.2, sme0=4000.0, ramda=0.2, kapa=0.04, ann=0.5, ocr=1.5, e=0.5)],
        2: [_layer(2, c=500.0, fai=30.0, e=1.8e6)],
        3: [_layer(3, zmf=1.2, sme0=4000.0, ramda=0.2, kapa=0.04, ann=0.5, bnn=0.5, beta=1.0, epss=0.6,
                   epsr=0.9, sitas=1.0, sitar=0.3, sd=1.0, sw=1.0, kse=1.0, c1=1.0, c2=1.0, c3=1.0, e=0.5)],
        4: [_layer(4, zmf=1.2, sme0=4000.0, ramda=0.2, kapa=0.04, mzm=1.0, azm=0.5, r0=1.0, bl=0.95, bz=1.0,
                   zet=0.0, e=0.5)],
        5: [_layer(5, c=300.0, fai=28.0, beta=5.0, e=1.5e6)],
        8: [_layer(8, c=0.7, eps0=0.8, sme0=4000.0, g0=120.0, zmf=1.2, ramda=0.1, etao=1.6, zet=0.5, ann=1.0,
                   bnn=1.0, damp1=1.0, damp2=0.1)],
    }
    w = default_fluid()
    w.update(dens=1000.0, vis=1e-3, vpu=30.0, press_type=0, acc_type=0, bhtype=0, pre_coe=10.0, strain_min=1.0)
    d.water = [dict(w), dict(w), dict(w)]
    parts = []
    for b, m in enumerate(models):
        parts.append((_lattice2d(0, nx, b * band_h, (b + 1) * band_h, dx), 2, m, 1 + (b % 3)))
    top = len(models) * band_h
    parts.append((_lattice2d(0, nx, top, top + water_h, dx), 1, 0))
    parts.append((_lattice2d(-3, nx + 3, -3, 0, dx), 0, 0))
    parts.append((_lattice2d(-3, 0, 0, top + water_h + 3, dx), 0, 0))
    parts.append((_lattice2d(nx, nx + 3, 0, top + water_h + 3, dx), 0, 0))
    return _assemble(d, parts, rng)


def moving_box_2d(if_move=1, nx=10, ny=8, dx=0.002, steps=12, seed=13, inip=2, speed=0.05) -> Deck:
    """2D Drucker-Prager soil block in a three-sided box whose driven part follows the
    Para_Boundary velocity (CPP_11-xsph.cpp:58-142):
      if_move 1: the right wall (type 0, matype 2) moves inwards at `speed`;
      if_move 2: the two top soil rows (etype 5) are pushed down at `speed`;
      if_move 3: an elastic structure plate (type 4, etype 5) on top is pushed down at `speed`."""
    rng = np.random.default_rng(seed)
    move = (-speed, 0.0, 0.0) if if_move == 1 else (0.0, -speed, 0.0)
    d = Deck(ndim=2, dr=dx, dt=0.2 * dx / 600.0, loop=steps, fop=1, gravity=(0.0, -9.81, 0.0),
             art_vis=((0.0, 0.0), (0.1, 0.0), (0.0, 0.0)), inip=inip, flbndy=2, if_move=if_move, move_vx=move,
             name=f"moving_box_2d_m{if_move}")
    d.model_flags[2] = 1
    d.soil_layers[2] = [dp_layer(c=200.0, fai=30.0, e=1.8e6, poi=0.3, dens=2000.0)]
    d.structure = {"dens": 2500.0, "eps0": 0.0, "cop": 0.0, "e": 2.0e7, "poi": 0.25}
    parts = [(_lattice2d(0, nx, 0, ny, dx), 2, 2),
             (_lattice2d(-3, nx + 3, -3, 0, dx), 0, 0),
             (_lattice2d(-3, 0, 0, ny + 4, dx), 0, 0),
             (_lattice2d(nx, nx + 3, 0, ny + 4, dx), 0, 2 if if_move == 1 else 0)]
    if if_move == 3:
        parts.append((_lattice2d(0, nx, ny, ny + 2, dx), 4, 0))
    deck = _assemble(d, parts, rng)
    if if_move == 2:
        top = (deck.type == 2) & (deck.xp[:, 1] > (ny - 2) * dx)
        deck.etype[top] = 5
    elif if_move == 3:
        deck.etype[deck.type == 4] = 5
    return deck


def art_stress_block(ndim=3, nx=5, ny=4, nz=6, dx=0.004, steps=10, seed=17, art_stress=0.3, step_regu=2,
                     inip=2, stretch=150.0) -> Deck:
    """Soil block of an elastic lower half and a Drucker-Prager upper half in a box (2D: nx by nz,
    3D: nx by ny by nz), with the artificial stress (CPP_9-5-artificialstress.cpp), strain/stress
    regularisation every `step_regu` steps and initial geostress damping switched on. The soil
    starts with the horizontal stretching velocity `stretch` * (x - x_centre), so that tensile
    principal stresses, the only ones the artificial stress acts on, appear within a few steps."""
    rng = np.random.default_rng(seed)
    grav = (0.0, -9.81, 0.0) if ndim == 2 else (0.0, 0.0, -9.81)
    d = Deck(ndim=ndim, dr=dx, dt=0.2 * dx / 600.0, loop=steps, fop=1, gravity=grav,
             art_vis=((0.0, 0.0), (0.1, 0.0), (0.0, 0.0)), art_stre_coe=art_stress, step_regu=step_regu,
             inip=inip, flbndy=2, damp_coe=0.05, name=f"art_stress_block_{ndim}d")
    d.model_flags[0] = 1
    d.model_flags[2] = 1
    d.soil_layers = {0: [_layer(0, e=3.0e6)], 2: [_layer(2, c=400.0, fai=30.0, e=1.8e6)]}
    h = nz // 2
    if ndim == 2:
        parts = [(_lattice2d(0, nx, 0, h, dx), 2, 0), (_lattice2d(0, nx, h, nz, dx), 2, 2),
                 (_lattice2d(-3, nx + 3, -3, 0, dx), 0, 0), (_lattice2d(-3, 0, 0, nz + 3, dx), 0, 0),
                 (_lattice2d(nx, nx + 3, 0, nz + 3, dx), 0, 0)]
    else:
        parts = [(_lattice3d(0, nx, 0, ny, 0, h, dx), 2, 0), (_lattice3d(0, nx, 0, ny, h, nz, dx), 2, 2),
                 (_box_walls_3d(nx, ny, nz + 3, dx), 0, 0)]
    deck = _assemble(d, parts, rng)
    soil = deck.type == 2
    deck.vxp[soil, 0] = stretch * (deck.xp[soil, 0] - 0.5 * nx * dx)
    return deck


def ssi_block(ndim=2, ss_type=1, nx=8, ny=4, nz=5, dx=0.002, steps=10, seed=21, myu=0.3, zeta=None, gap=0.4,
              v0=(0.4, -0.6, 0.3), acc_type=1, water=0) -> Deck:
    """Soil-structure interaction (Para_SSInter type 1 rigid, 2 deformable, 3 penalty;
    CPP_10-1-interact_ss.cpp): a Drucker-Prager soil block (2D: nx by nz, 3D: nx by ny by nz)
    moving at v0 onto an elastic structure plate (type 4, two rows / layers, 2 dx wider than the
    block on each side) that rests on a 3-layer rigid floor. The block's lowest row sits `gap` dx
    above the plate's top row, so soil particles are in contact from the first step; v0 has a
    tangential part so that the Coulomb cap (friction `myu`) acts. zeta defaults to 0.5, and to 0.999
    for the penalty form, whose normal force (1 - zeta) 2 m (dr - dp) / dt^2 is otherwise explosive
    at this time step. acc_type 2 switches on coupled_Water_Soil_Structure; `water` > 0 adds that
    many lattice columns of water (type 1) beside the block, on the plate."""
    if zeta is None:
        zeta = 0.999 if ss_type == 3 else 0.5
    rng = np.random.default_rng(seed)
    grav = (0.0, -9.81, 0.0) if ndim == 2 else (0.0, 0.0, -9.81)
    d = Deck(ndim=ndim, dr=dx, dt=0.2 * dx / 600.0, loop=steps, fop=1, gravity=grav, ss_type=ss_type, ss_myu=myu,
             ss_zeta=zeta, ss_acc_type=acc_type, art_vis=((0.0, 0.0), (0.1, 0.0), (0.0, 0.0)), flbndy=2,
             name=f"ssi_block_{ndim}d_t{ss_type}")
    if water:
        w = default_fluid()
        w.update(dens=1000.0, vis=1e-3, vpu=30.0, press_type=0, acc_type=0, bhtype=0, pre_coe=1.0, strain_min=1.0)
        d.water = [dict(w), dict(w), dict(w)]
    d.model_flags[2] = 1
    # with water the soil's void ratio sets the fluid porosity of the water EOS (0 would divide by 0)
    d.soil_layers[2] = [dp_layer(c=300.0, fai=30.0, e=1.8e6, poi=0.3, dens=2000.0, eps0=0.6 if water else 0.0)]
    d.structure = {"dens": 2500.0, "eps0": 0.0, "cop": 0.0, "e": 2.0e7, "poi": 0.25}
    drop = (1.0 - gap) * dx
    if ndim == 2:
        soil = _lattice2d(0, nx, 2, 2 + nz, dx)
        soil[:, 1] -= drop
        parts = [(soil, 2, 2), (_lattice2d(-2, nx + water + 2, 0, 2, dx), 4, 0),
                 (_lattice2d(-5, nx + water + 5, -3, 0, dx), 0, 0)]
        if water:
            parts.append((_lattice2d(nx, nx + water, 2, 2 + nz, dx), 1, 0))
    else:
        soil = _lattice3d(0, nx, 0, ny, 2, 2 + nz, dx)
        soil[:, 2] -= drop
        parts = [(soil, 2, 2), (_lattice3d(-2, nx + water + 2, -2, ny + 2, 0, 2, dx), 4, 0),
                 (_lattice3d(-5, nx + water + 5, -5, ny + 5, -3, 0, dx), 0, 0)]
        if water:
            parts.append((_lattice3d(nx, nx + water, 0, ny, 2, 2 + nz, dx), 1, 0))
    deck = _assemble(d, parts, rng)
    s = deck.type == 2
    # v0 = (tangential x, normal towards the plate, second tangential (3D))
    deck.vxp[s] = (v0[0], v0[1], 0.0) if ndim == 2 else (v0[0], v0[2], v0[1])
    return deck


def tran_box(ndim=2, flbndy=3, nx=8, ny=3, nz_soil=4, nz_water=3, dx=0.002, steps=8, seed=31, v0=0.2,
             gap=None, myu=0.3, zeta=0.5, vdown=0.0) -> Deck:
    """Soil under water in a box whose walls use the boundary treatment of Tran et al.
    (flbndy 3 no slip, 4 free slip, 13 / 14 the same after velocity_set_domain;
    CPP_13-boundary.cpp:711-1237, CPP_9-4-derivation_stress.cpp:725-820). Soil (DP, void ratio
    0.6) and water start sliding at v0 along the floor, so that the wall velocities and stresses
    each phase sees differ from zero. flbndy 5 (free-solid walls, CPP_13-boundary.cpp:1239-1515,
    CPP_9-4-derivation_stress.cpp:900-990) takes the wall friction myu and factor zeta from
    Para_SSInter; `gap` moves the soil and water down so that their lowest row sits gap dx above
    the floor's top row (contact needs less than 0.5 dx); vdown adds a velocity towards the floor."""
    rng = np.random.default_rng(seed)
    grav = (0.0, -9.81, 0.0) if ndim == 2 else (0.0, 0.0, -9.81)
    d = Deck(ndim=ndim, dr=dx, dt=0.2 * dx / 600.0, loop=steps, fop=1, gravity=grav, flbndy=flbndy, coe_bndy=0.1,
             art_vis=((0.1, 0.0), (0.1, 0.0), (0.0, 0.0)), name=f"tran_box_{ndim}d_b{flbndy}")
    if flbndy == 5:
        d.ss_myu, d.ss_zeta = myu, zeta
    d.model_flags[2] = 1
    d.soil_layers[2] = [dp_layer(c=300.0, fai=30.0, e=1.8e6, poi=0.3, dens=2000.0, eps0=0.6)]
    w = default_fluid()
    w.update(dens=1000.0, vis=1e-3, vpu=30.0, press_type=0, acc_type=0, bhtype=0, pre_coe=1.0, strain_min=1.0)
    d.water = [dict(w), dict(w), dict(w)]
    top = nz_soil + nz_water
    if ndim == 2:
        parts = [(_lattice2d(0, nx, 0, nz_soil, dx), 2, 2), (_lattice2d(0, nx, nz_soil, top, dx), 1, 0),
                 (_lattice2d(-3, nx + 3, -3, 0, dx), 0, 0), (_lattice2d(-3, 0, 0, top + 3, dx), 0, 0),
                 (_lattice2d(nx, nx + 3, 0, top + 3, dx), 0, 0)]
    else:
        parts = [(_lattice3d(0, nx, 0, ny, 0, nz_soil, dx), 2, 2), (_lattice3d(0, nx, 0, ny, nz_soil, top, dx), 1, 0),
                 (_box_walls_3d(nx, ny, top + 3, dx), 0, 0)]
    deck = _assemble(d, parts, rng)
    moving = (deck.type == 1) | (deck.type == 2)
    if gap is not None:
        deck.xp[moving, ndim - 1] -= (1.0 - gap) * dx
    deck.vxp[moving, 0] = v0
    deck.vxp[moving, ndim - 1] = -vdown
    return deck


def drag_box(ndim=2, fli=2, nx=8, ny=3, nz_soil=4, nz_water=3, dx=0.002, steps=6, seed=51, eps0=0.6, ds=4e-3,
             vis=1e-3, v0=0.3, c_water=30.0) -> Deck:
    """Soil under water in a rigid-walled box (flbndy 2) with the fluid-solid drag of Tsuji et al.
    (fli 2: Ergun below porosity 0.8, Wen-Yu above; CPP_10-interact1.cpp:115-267) or of Peng et al.
    (fli 3: Kozeny-Carman, turbulent term above Reynolds number 60; :270-420). Soil particles
    start at porosity eps0 / (1 + eps0) with grain size ds; every water particle slides along x at
    its own speed, uniform in [0, 2 v0), over soil sliding at -v0 / 2, so the pairs' Reynolds
    numbers spread over the branches of the drag law."""
    rng = np.random.default_rng(seed)
    grav = (0.0, -9.81, 0.0) if ndim == 2 else (0.0, 0.0, -9.81)
    d = Deck(ndim=ndim, dr=dx, dt=0.2 * dx / 600.0, loop=steps, fop=1, gravity=grav, fli=fli, interc=1.0, flbndy=2,
             art_vis=((0.1, 0.0), (0.1, 0.0), (0.0, 0.0)), name=f"drag_box_{ndim}d_fli{fli}")
    d.model_flags[2] = 1
    layer = dp_layer(c=1000.0, fai=25.0, e=1.8e6, poi=0.3, dens=2650.0, eps0=eps0, cop=1e-4)
    layer["ds"] = ds
    d.soil_layers[2] = [layer]
    w = default_fluid()
    por = eps0 / (1 + eps0)
    B = 1000.0 * c_water ** 2 / 7.0
    w.update(dens=1000.0, vis=vis, vpu=c_water, press_type=0, acc_type=0, bhtype=0,
             pre_coe=B / (por * 1000.0 * 9.81 * dx), strain_min=1.0)
    d.water = [dict(w), dict(w), dict(w)]
    top = nz_soil + nz_water
    if ndim == 2:
        parts = [(_lattice2d(0, nx, 0, nz_soil, dx), 2, 2), (_lattice2d(0, nx, nz_soil, top, dx), 1, 0),
                 (_lattice2d(-3, nx + 3, -3, 0, dx), 0, 0), (_lattice2d(-3, 0, 0, top + 3, dx), 0, 0),
                 (_lattice2d(nx, nx + 3, 0, top + 3, dx), 0, 0)]
    else:
        parts = [(_lattice3d(0, nx, 0, ny, 0, nz_soil, dx), 2, 2), (_lattice3d(0, nx, 0, ny, nz_soil, top, dx), 1, 0),
                 (_box_walls_3d(nx, ny, top + 3, dx), 0, 0)]
    deck = _assemble(d, parts, rng)
    water = deck.type == 1
    deck.vxp[water, 0] = v0 * rng.uniform(0.0, 2.0, int(water.sum()))
    deck.vxp[deck.type == 2, 0] = -0.5 * v0
    return deck


def air_water_box(ndim=2, acc_type=0, nx=8, ny=3, nz_soil=3, nz_fluid=3, dx=0.002, steps=6, seed=61,
                  eps0=0.6, perm=1e-4, v0=0.2, paa0=200.0) -> Deck:
    """Three phases in a rigid-walled box (flbndy 2): unsaturated Cam-Clay soil (matype 3, void
    ratio eps0, saturation starting at sitar) under water (type 1) on the left half and air
    (type 3, fla 1) on the right half, with Darcy drag (fli 1, permtype 1) for both fluids
    (CPP_10-interact1.cpp:15-110; the air term, x2000, scales with (1 - satu)^2 and so needs an
    unsaturated soil). acc_type selects the fluids' momentum sums for water and air together
    (0: fluid_acceleration_water / _air, 1: the _minorwater / _minorair forms,
    CPP_9-4-derivation_stress.cpp:12-351). The left wall's particles are matype 1, the walls whose
    pressure follows the Tait equation against their own rhop (CPP_7-1-eos_CR.cpp:37-41); air takes
    the reference's equation of state with its comma operator (:31-35) from its initial pressure
    paa0 (press_type 0). Fluids slide along x at v0 and the soil at -v0/2, so drag and viscosity
    act from the first step."""
    rng = np.random.default_rng(seed)
    grav = (0.0, -9.81, 0.0) if ndim == 2 else (0.0, 0.0, -9.81)
    d = Deck(ndim=ndim, dr=dx, dt=0.2 * dx / 600.0, loop=steps, fop=1, gravity=grav, fla=1, fli=1, permtype=1,
             pers=perm, perl=perm, perm=perm, interc=1.0, flbndy=2,
             art_vis=((0.1, 0.0), (0.1, 0.0), (0.1, 0.0)), name=f"air_water_box_{ndim}d_a{acc_type}")
    d.model_flags[3] = 1
    d.soil_layers[3] = [_layer(3, zmf=1.2, sme0=4000.0, ramda=0.2, kapa=0.04, ann=0.5, bnn=0.5, beta=1.0, epss=0.6,
                               epsr=0.9, sitas=1.0, sitar=0.3, sd=1.0, sw=1.0, kse=1.0, c1=1.0, c2=1.0, c3=1.0,
                               e=0.5, eps0=eps0, cop=perm)]
    w = default_fluid()
    w.update(dens=1000.0, vis=1e-3, vpu=30.0, press_type=0, acc_type=acc_type, bhtype=0, pre_coe=1.0,
             strain_min=1.0)
    # the deck file carries acc_type for the first water material only (CPP_2-initialf.cpp:417-420)
    d.water = [dict(w), dict(w, acc_type=0), dict(w, acc_type=0)]
    a = default_fluid()
    a.update(dens=1.293, paa0=paa0, vis=1.8e-5, vpu=30.0, press_type=0, acc_type=acc_type, pre_coe=1.0)
    d.air = a
    top, hx = nz_soil + nz_fluid, nx // 2
    if ndim == 2:
        parts = [(_lattice2d(0, nx, 0, nz_soil, dx), 2, 3), (_lattice2d(0, hx, nz_soil, top, dx), 1, 0),
                 (_lattice2d(hx, nx, nz_soil, top, dx), 3, 0),
                 (_lattice2d(-3, nx + 3, -3, 0, dx), 0, 0), (_lattice2d(-3, 0, 0, top + 3, dx), 0, 1),
                 (_lattice2d(nx, nx + 3, 0, top + 3, dx), 0, 0)]
    else:
        parts = [(_lattice3d(0, nx, 0, ny, 0, nz_soil, dx), 2, 3), (_lattice3d(0, hx, 0, ny, nz_soil, top, dx), 1, 0),
                 (_lattice3d(hx, nx, 0, ny, nz_soil, top, dx), 3, 0),
                 (_lattice3d(-3, nx + 3, -3, ny + 3, -3, 0, dx), 0, 0),
                 (_lattice3d(-3, 0, -3, ny + 3, 0, top + 3, dx), 0, 1),
                 (_lattice3d(nx, nx + 3, -3, ny + 3, 0, top + 3, dx), 0, 0),
                 (_lattice3d(0, nx, -3, 0, 0, top + 3, dx), 0, 0),
                 (_lattice3d(0, nx, ny, ny + 3, 0, top + 3, dx), 0, 0)]
    deck = _assemble(d, parts, rng)
    fluid = (deck.type == 1) | (deck.type == 3)
    deck.vxp[fluid, 0] = v0
    deck.vxp[deck.type == 2, 0] = -0.5 * v0
    return deck


def slope_pile_3d(nx=200, ny=50, nz=100, dx=0.1, model=4, steps=1000, seed=71, pile=(10, 12), pile_x=0.6,
                  cut=(0.5, 0.75, 0.1), pile_shift=(0.6, 0.0, 0.0), ss_type=2, myu=0.3, zeta=0.5, walls=3,
                  v0=(0.0, 0.0, 0.0)) -> Deck:
    """3D slope with an embedded pile (soil-structure interaction). Soil on an nx by ny by nz
    lattice (x along the slope, y across, z up) with a 1:2 (vertical:horizontal) cut: the surface
    is at nz for x below cut[0] nx, falls by one cell per two along x, and stays at cut[1] nz past
    the face; soil is modified Cam-Clay, model 4 (AMCC, the reference's no-op tangent: lambda 0.2,
    kappa 0.04, M 1.2, OCR 1.5, e0 0.8) or model 1 (SB sub-loading Cam-Clay with the same
    constants). The pile is a pile[0] by pile[1] column of elastic structure particles
    (type 4, E = 30 GPa, nu 0.2, rho 2500) centred across the slope at pile_x nx, from the base to
    cut[2] nz above the surface there (below it for cut[2] < 0), in place of the soil it displaces,
    shifted by pile_shift dx (by default 0.6 dx downslope, so that its downslope face sits 0.4 dx
    from the soil and the contact acts from the first step); the soil-pile contact is Para_SSInter
    type ss_type
    (2: deformable, CPP_10-1-interact_ss.cpp:352-691). Rigid base and four side walls of `walls`
    layers; seeded 1e-6 dx jitter; the soil starts at velocity v0 (the contact force is
    m v_rel / dt, so a soil at rest against a pile at rest feels none). The defaults
    (200 x 50 x 100 at dx = 0.1) give about 0.9 M soil and 12 k pile particles."""
    rng = np.random.default_rng(seed)
    # time step from the pile's elastic wave speed sqrt(E / rho) = 3.5 km/s (the soils' 600 m/s
    # would let the pile's stress waves grow without bound)
    c_pile = float(np.sqrt(3.0e10 / 2500.0))
    d = Deck(ndim=3, dr=dx, dt=0.2 * dx / c_pile, loop=steps, fop=steps, gravity=(0.0, 0.0, -9.81), flbndy=2,
             ss_type=ss_type, ss_myu=myu, ss_zeta=zeta, ss_acc_type=1,
             art_vis=((0.0, 0.0), (0.1, 0.0), (0.0, 0.0)), name=f"slope_pile3d_m{model}_{nx}x{ny}x{nz}")
    d.model_flags[model] = 1
    cam = dict(poi=0.3, dens=2000.0, eps0=0.8, zmf=1.2, sme0=20000.0, fs=1.0, ramda=0.2, kapa=0.04, cop=1e-4,
               e=0.8, ds=1.0, ocr=1.5, ann=0.5, mzm=1.0, azm=0.5, r0=1.0, bl=0.95, bz=1.0, zet=0.0)
    d.soil_layers[model] = [_layer(model, **cam)]
    d.structure = {"dens": 2500.0, "eps0": 0.0, "cop": 0.0, "e": 3.0e10, "poi": 0.2}
    ix = np.arange(nx)
    x0, ztoe = cut[0] * nx, int(round(cut[1] * nz))
    surf = np.where(ix < x0, nz, np.maximum(ztoe, nz - np.floor((ix - x0) / 2.0).astype(np.int64)))
    px0 = int(round(pile_x * nx)) - pile[0] // 2
    py0 = ny // 2 - pile[1] // 2
    ptop = int(surf[px0 + pile[0] // 2]) + int(round(cut[2] * nz))
    inpile = np.zeros((nx, ny), bool)
    inpile[px0:px0 + pile[0], py0:py0 + pile[1]] = True
    cols = []
    for i in range(nx):
        for j in range(ny):
            if not inpile[i, j]:
                cols.append((i, j, int(surf[i])))
    ci = np.array([c[0] for c in cols]); cj = np.array([c[1] for c in cols]); ch = np.array([c[2] for c in cols])
    I = np.repeat(ci, ch); J = np.repeat(cj, ch)
    K = np.concatenate([np.arange(h) for h in ch])
    soil = (np.stack([I, J, K], axis=1) + 0.5) * dx
    pile_p = _lattice3d(px0, px0 + pile[0], py0, py0 + pile[1], 0, ptop, dx)
    pile_p += np.asarray(pile_shift, dtype=np.float64) * dx
    hw = nz + walls
    wl = walls
    parts = [(soil, 2, model), (pile_p, 4, 0),
             (_lattice3d(-wl, nx + wl, -wl, ny + wl, -wl, 0, dx), 0, 0),
             (_lattice3d(-wl, 0, -wl, ny + wl, 0, hw, dx), 0, 0),
             (_lattice3d(nx, nx + wl, -wl, ny + wl, 0, hw, dx), 0, 0),
             (_lattice3d(0, nx, -wl, 0, 0, hw, dx), 0, 0),
             (_lattice3d(0, nx, ny, ny + wl, 0, hw, dx), 0, 0)]
    deck = _assemble(d, parts, rng)
    deck.vxp[deck.type == 2] = v0
    return deck


def vibration_box(ndim=2, flv=1, nx=8, ny=3, nz=5, dx=0.002, steps=6, seed=81, amp=None, v0=0.05,
                  damp=(50.0, 1e-3)) -> Deck:
    """Soil on a shaking base (flv 1: the base walls, matype 3, and the side walls, matype 5, take
    the acceleration of the vibration table; flv 2: the base, matype 4, and the sides take its
    velocity, CPP_9-1-acceleration_vibration.cpp). The sides are free-field boundaries: a column of
    free-field soil (matype 52, its normal x as permtype 1) next to each side wall (matype 5,
    normal x; in 3D the y walls take normal y), CPP_19-FreeFieldParticles*.cpp. The soil (DP with
    damp1 / damp2 of soil_damping_vibration) starts moving at v0 along x. The table has eight loads
    per period of 4 dt and applies from t = 1.5 dt to (steps - 0.5) dt; its sixth load exceeds 1000
    in magnitude, which zeroes the walls' motion for that step."""
    rng = np.random.default_rng(seed)
    grav = (0.0, -9.81, 0.0) if ndim == 2 else (0.0, 0.0, -9.81)
    d = Deck(ndim=ndim, dr=dx, dt=0.2 * dx / 600.0, loop=steps, fop=1, gravity=grav, flbndy=2, coe_bndy=0.1,
             flv=flv, art_vis=((0.1, 0.0), (0.1, 0.0), (0.0, 0.0)), name=f"vibration_{ndim}d_v{flv}")
    d.model_flags[2] = 1
    lay = dp_layer(c=500.0, fai=30.0, e=1.8e6, poi=0.3, dens=2000.0, eps0=0.6)
    lay.update(damp1=damp[0], damp2=damp[1])
    d.soil_layers[2] = [lay]
    base = 3 if flv == 1 else 4
    top = nz + 2
    if ndim == 2:
        parts = [(_lattice2d(1, nx - 1, 0, nz, dx), 2, 2), (_lattice2d(0, 1, 0, nz, dx), 2, 52, 1),
                 (_lattice2d(nx - 1, nx, 0, nz, dx), 2, 52, 1), (_lattice2d(-3, nx + 3, -3, 0, dx), 0, base),
                 (_lattice2d(-3, 0, 0, top, dx), 0, 5, 1), (_lattice2d(nx, nx + 3, 0, top, dx), 0, 5, 1)]
    else:
        parts = [(_lattice3d(1, nx - 1, 0, ny, 0, nz, dx), 2, 2), (_lattice3d(0, 1, 0, ny, 0, nz, dx), 2, 52, 1),
                 (_lattice3d(nx - 1, nx, 0, ny, 0, nz, dx), 2, 52, 1),
                 (_lattice3d(-3, nx + 3, -3, ny + 3, -3, 0, dx), 0, base),
                 (_lattice3d(-3, 0, -3, ny + 3, 0, top, dx), 0, 5, 1),
                 (_lattice3d(nx, nx + 3, -3, ny + 3, 0, top, dx), 0, 5, 1),
                 (_lattice3d(0, nx, -3, 0, 0, top, dx), 0, 5, 2), (_lattice3d(0, nx, ny, ny + 3, 0, top, dx), 0, 5, 2)]
    deck = _assemble(d, parts, rng)
    soil = deck.type == 2
    deck.vxp[soil, 0] = v0
    if amp is None:
        amp = 5.0 if flv == 1 else 0.05
    ph = np.arange(8) * (2.0 * math.pi / 8.0)
    loads = np.zeros((8, 3))
    loads[:, 0] = amp * np.sin(ph)
    loads[:, ndim - 1] = 0.5 * amp * np.cos(ph)
    loads[5, 0] = 2000.0
    deck.vib_ttime, deck.vib_nsteps = 4.0 * d.dt, 8
    deck.vib_sttime, deck.vib_edtime = 1.5 * d.dt, (steps - 0.5) * d.dt
    deck.vib_loads = loads
    return deck


def rain_box(ndim=2, rtype=1, nx=8, ny=3, nz=4, dx=0.002, steps=6, seed=91, rows=2, vdrop=-0.5) -> Deck:
    """Drucker-Prager soil in a rigid-walled box (flbndy 2) with the dynamic water sets of flr 1
    (CPP_15-1-rain.cpp, CPP_15-2-partidel.cpp). rtype 1: `rows` rows of reserve particles (type 7,
    etype 1) above the soil, last in the particle array, fall in as rain one row per period; rtype
    4: two groups of reserve particles (matype 1 then 2) are injected radially from their centre as
    cement, a group per period; rtype 2 / 3: a water layer of etype 7 over the soil loses its
    left (right) columns back to the reserve."""
    rng = np.random.default_rng(seed)
    grav = (0.0, -9.81, 0.0) if ndim == 2 else (0.0, 0.0, -9.81)
    d = Deck(ndim=ndim, dr=dx, dt=0.2 * dx / 600.0, loop=steps, fop=1, gravity=grav, flbndy=2, coe_bndy=0.1, flr=1,
             art_vis=((0.1, 0.0), (0.1, 0.0), (0.0, 0.0)), name=f"rain_{ndim}d_r{rtype}")
    d.model_flags[2] = 1
    d.soil_layers[2] = [dp_layer(c=300.0, fai=30.0, e=1.8e6, poi=0.3, dens=2000.0, eps0=0.6)]
    w = default_fluid()
    w.update(dens=1000.0, vis=1e-3, vpu=30.0, press_type=0, acc_type=0, bhtype=0, pre_coe=1.0, strain_min=1.0,
             porosity=1.0)
    w1, w2 = dict(w), dict(w)
    w1.update(dens=1200.0, porosity=0.9)
    w2.update(dens=1400.0, porosity=0.8)
    d.water = [w, w1, w2]
    H = nz + 10
    lat = (lambda a, b, c_, e: _lattice2d(a, b, c_, e, dx)) if ndim == 2 else \
        (lambda a, b, c_, e: _lattice3d(a, b, 0, ny, c_, e, dx))
    parts = [(lat(0, nx, 0, nz), 2, 2)]
    if rtype in (2, 3):
        parts.append((lat(0, nx, nz, nz + 3), 1, 0))
    if ndim == 2:
        walls = [(_lattice2d(-3, nx + 3, -3, 0, dx), 0, 0), (_lattice2d(-3, 0, 0, H + 3, dx), 0, 0),
                 (_lattice2d(nx, nx + 3, 0, H + 3, dx), 0, 0)]
    else:
        walls = [(_box_walls_3d(nx, ny, H + 3, dx), 0, 0)]
    parts += walls
    reserve = []
    if rtype == 1:
        for r in range(rows):
            reserve.append((lat(0, nx, H - 1 - 2 * r, H - 2 * r), 7, 0))
    elif rtype == 4:
        reserve.append((lat(2, nx - 2, H - 4, H - 2), 7, 1))
        reserve.append((lat(2, nx - 2, H - 2, H), 7, 2))
    parts += reserve
    deck = _assemble(d, parts, rng)
    if rtype in (2, 3):
        deck.etype[deck.type == 1] = 7
    nres = int(sum(p[0].shape[0] for p in reserve))
    nhor = nx if ndim == 2 else nx * ny
    if rtype == 4:
        nhor = nres // 2
    deck.rain = (rtype, nres, nhor, 1.0, vdrop)
    return deck


def eob_box(ndim=2, nx=8, ny=3, nz=6, dx=0.002, steps=6, seed=93) -> Deck:
    """Excavation and backfill (fleob 1, CPP_16-EOB.cpp): a Drucker-Prager block in a rigid-walled
    box whose top-right part is excavated at step 2 (type 2 -> 7) while a backfill next to it (type
    7, etype 2) turns into soil at step 4."""
    rng = np.random.default_rng(seed)
    grav = (0.0, -9.81, 0.0) if ndim == 2 else (0.0, 0.0, -9.81)
    d = Deck(ndim=ndim, dr=dx, dt=0.2 * dx / 600.0, loop=steps, fop=1, gravity=grav, flbndy=2, coe_bndy=0.1,
             fleob=1, art_vis=((0.1, 0.0), (0.1, 0.0), (0.0, 0.0)), name=f"eob_{ndim}d")
    d.model_flags[2] = 1
    d.soil_layers[2] = [dp_layer(c=300.0, fai=30.0, e=1.8e6, poi=0.3, dens=2000.0, eps0=0.6)]
    lat = (lambda a, b, c_, e: _lattice2d(a, b, c_, e, dx)) if ndim == 2 else \
        (lambda a, b, c_, e: _lattice3d(a, b, 0, ny, c_, e, dx))
    h = nx // 2
    parts = [(lat(0, nx, 0, nz - 2), 2, 2), (lat(h, nx, nz - 2, nz), 2, 2), (lat(0, h, nz - 2, nz), 7, 2)]
    if ndim == 2:
        parts += [(_lattice2d(-3, nx + 3, -3, 0, dx), 0, 0), (_lattice2d(-3, 0, 0, nz + 3, dx), 0, 0),
                  (_lattice2d(nx, nx + 3, 0, nz + 3, dx), 0, 0)]
    else:
        parts += [(_box_walls_3d(nx, ny, nz + 3, dx), 0, 0)]
    deck = _assemble(d, parts, rng)
    n0 = parts[0][0].shape[0]
    n1 = parts[1][0].shape[0]
    n2 = parts[2][0].shape[0]
    deck.etype[n0:n0 + n1] = 7      # excavated: swaps to type 7
    deck.etype[n0 + n1:n0 + n1 + n2] = 2  # backfill: swaps to soil
    deck.ecp = [(n0 + 1, n0 + n1, 2), (n0 + n1 + 1, n0 + n1 + n2, 4)]
    return deck


def slope_2d(height=5.0, dx=0.1, steps=3000, seed=0, crest=1.5, toe=1.5, base=4.0, walls=3) -> Deck:
    """2D soil slope for the random-field batch: a crest of `crest`*H, a 1:2 (vertical:horizontal)
    face, a toe of `toe`*H and a `base` m layer under all of it, Drucker-Prager soil (mean
    c = 10 kPa, phi = 25, E = 20 MPa), rigid floor and side walls of `walls` layers, lattice at
    (i + 1/2) dx with seeded jitter. H = 5..15 m at dx = 0.1 gives about 16k..86k particles."""
    rng = np.random.default_rng(seed)
    H = float(height)
    d = Deck(ndim=2, dr=dx, dt=0.2 * dx / 600.0, loop=steps, fop=steps, gravity=(0.0, -9.81, 0.0),
             art_vis=((0.0, 0.0), (0.1, 0.0), (0.0, 0.0)), flbndy=2, name=f"slope2d_H{H:g}")
    d.model_flags[2] = 1
    d.soil_layers[2] = [dp_layer(c=10000.0, fai=25.0, e=2.0e7, poi=0.3, dens=2000.0)]
    L = (crest + 2.0 + toe) * H
    nx = int(round(L / dx))
    nb = int(round(base / dx))
    xs = (np.arange(nx) + 0.5) * dx
    top = np.where(xs < crest * H, H, np.where(xs < (crest + 2.0) * H, H - (xs - crest * H) / 2.0, 0.0))
    ny = nb + np.floor(top / dx).astype(np.int64)
    ix = np.repeat(np.arange(nx), ny)
    iy = np.concatenate([np.arange(k) for k in ny])
    soil = np.zeros((ix.size, 3))
    soil[:, 0] = (ix + 0.5) * dx
    soil[:, 1] = (iy + 0.5) * dx
    hw = nb + int(np.ceil(H / dx)) + 2
    parts = [(soil, 2, 2), (_lattice2d(-walls, nx + walls, -walls, 0, dx), 0, 0),
             (_lattice2d(-walls, 0, 0, hw, dx), 0, 0), (_lattice2d(nx, nx + walls, 0, nb + 2, dx), 0, 0)]
    return _assemble(d, parts, rng)
