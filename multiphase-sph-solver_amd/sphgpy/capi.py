"""ctypes view of include/sphg.h: parameter structs, the product library and a thin context wrapper.

The structs mirror the reference parameter classes (Class_Problem.h, Class_ParticleProperty.h)
field for field; ``problem_from_deck``/``materials_from_deck`` fill them the way the reference's
deck reader does (CPP_2-initialf.cpp:22-1018).
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Dict, Optional

import numpy as np

from .deck import Deck, SOIL_FIELDS, FLUID_FIELDS, STRUCT_FIELDS

_D = C.c_double
_I = C.c_int

PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REPO_DIR = os.path.dirname(PKG_DIR)
LIB_PATH = os.path.join(PKG_DIR, "_build", "libsphg.so")
SOLVE_PATH = os.path.join(PKG_DIR, "_build", "sphg_solve")  # project runner (sphg_solve_project)
MAX_NB = 99


class Problem(C.Structure):
    _fields_ = ([(n, _I) for n in "fla fls flr fli flv fleob flts flspv flknl ntd flbndy".split()] +
                [(n, _I) for n in "ndim fop inip loop ntotal step_regu dttype".split()] +
                [("dr", _D), ("dt", _D), ("tt", _D), ("xsph", _D * 3), ("art_vis", (_D * 2) * 3),
                 ("art_stre_coe", _D), ("threshold_min", _D), ("threshold_max", _D),
                 ("gx", _D), ("gy", _D), ("gz", _D), ("if_cal", _I), ("if_move", _I),
                 ("coe_bndy", _D), ("move_vx", _D * 3), ("interc", _D), ("permtype", _I),
                 ("pers", _D), ("perl", _D), ("perm", _D), ("ss_type", _I), ("ss_myu", _D),
                 ("ss_zeta", _D), ("ss_acc_type", _I)] +
                [(n, _I) for n in "flagem flagsb flagdp flagum flagam flagbui flagpl flagve flagfocc flagep".split()] +
                [("alphats", _D)] +
                [(n, _D) for n in "xmin xmax ymin ymax zmin zmax".split()] +
                [("soil_min", _D * 3), ("soil_max", _D * 3), ("water_min", _D * 3), ("water_max", _D * 3)])


class Soil(C.Structure):
    _fields_ = [(n, _D) for n in SOIL_FIELDS]


class Fluid(C.Structure):
    _fields_ = [(n, _D) for n in "vis vpu dens paa0 porosity fai c".split()] + \
               [(n, _I) for n in "bhtype press_type acc_type".split()] + \
               [("pre_coe", _D), ("strain_min", _D)]


class Structure(C.Structure):
    _fields_ = [(n, _D) for n in STRUCT_FIELDS]


class Materials(C.Structure):
    _fields_ = [("soil", Soil * 30), ("water", Fluid * 3), ("air", Fluid), ("structure", Structure)]


KL_TERMS = 200


class Box(C.Structure):
    _fields_ = [(n, _D) for n in "xmin xmax ymin ymax zmin zmax".split()]


class KLField(C.Structure):
    """sphg_kl_field: one spatially variable property and its K-L modes."""
    _fields_ = [("ndim", _I), ("dist", _I), ("field", _I), ("mean_value", _D), ("sd_value", _D),
                ("coe_depth", _D), ("rl", _D * 3), ("a", _D * 3), ("x_trans", _D * 3), ("soil_max", _D),
                ("water_max", _D), ("wi", (_D * 3) * KL_TERMS), ("eigval", _D * KL_TERMS),
                ("order", (_I * 3) * KL_TERMS)]

    def modes(self):
        """(wi[200][3], eigval[200], order[200][3]) as arrays."""
        wi = np.ctypeslib.as_array(self.wi).reshape(KL_TERMS, 3).copy()
        ev = np.ctypeslib.as_array(self.eigval).copy()
        od = np.ctypeslib.as_array(self.order).reshape(KL_TERMS, 3).copy()
        return wi, ev, od


class RfField(C.Structure):
    """sphg_rf_field: one property of a correlation-matrix random field."""
    _fields_ = [("ndim", _I), ("dist", _I), ("field", _I), ("acf", _I), ("mean_value", _D), ("sd_value", _D),
                ("coe_depth", _D), ("rl", _D * 3), ("beta", _D), ("beta_1", _D), ("cell_ndx", _D), ("dr", _D),
                ("box_nb", _D * 6), ("soil_max", _D), ("water_max", _D), ("nc", _I * 3), ("cell_total", _I),
                ("base", _D * 3), ("cell_dx", _D)]


def rf_field(deck: Deck, dist: int, mean: float, sd: float, rl, field: int = 1, acf: int = 1, cell_ndx: float = 4.0,
             coe_depth: float = 0.0, beta: float = 0.0, beta_1: float = 0.0) -> RfField:
    """RfField of a deck: the non-boundary box and phase maxima as input_par forms them
    (CPP_2-initialf.cpp:911-1017) on the vertical axis (y in 2D, z in 3D)."""
    f = RfField()
    f.ndim, f.dist, f.field, f.acf = deck.ndim, dist, field, acf
    f.mean_value, f.sd_value, f.coe_depth = mean, sd, coe_depth
    for k in range(3):
        f.rl[k] = rl[k] if k < len(rl) else 0.0
    f.beta, f.beta_1, f.cell_ndx, f.dr = beta, beta_1, cell_ndx, deck.dr
    nb = (deck.type != 0) & (deck.type != 7)
    box = []
    for k in range(3):
        mx = max(-10000.0, float(deck.xp[nb, k].max())) if nb.any() else -10000.0
        mn = min(10000.0, float(deck.xp[nb, k].min())) if nb.any() else 10000.0
        box += [mn, mx]
    for k in range(6):
        f.box_nb[k] = box[k]
    _, _, ph = deck.domain()
    v = 1 if deck.ndim == 2 else 2
    f.soil_max = float(ph["soil"][1][v])
    f.water_max = float(ph["water"][1][v])
    return f


def rf_cells(f: RfField, xp: np.ndarray, ty: np.ndarray):
    """sphg_rf_cells (host): (rc, cell_of[n]); fills f's cell partition."""
    xp = np.ascontiguousarray(xp, dtype=np.float64)
    ty = np.ascontiguousarray(ty, np.int32)
    cell_of = np.zeros(xp.shape[0], np.int32)
    rc = product_lib().lib.sphg_rf_cells(C.byref(f), xp.shape[0], _dptr(xp), _iptr(ty), _iptr(cell_of))
    return int(rc), cell_of


def rf_generate(f: RfField, xp: np.ndarray, ty: np.ndarray, zeta: np.ndarray, fill: float = 0.0,
                want_l: bool = False):
    """sphg_rf_generate: (rc, value[nreal][n], L or None, factor_ms); zeta is (nreal, cell_total)."""
    xp = np.ascontiguousarray(xp, dtype=np.float64)
    ty = np.ascontiguousarray(ty, np.int32)
    zeta = np.ascontiguousarray(np.atleast_2d(zeta), dtype=np.float64)
    nreal = zeta.shape[0]
    value = np.zeros((nreal, xp.shape[0]))
    g = RfField.from_buffer_copy(f)
    rc0, _ = rf_cells(g, xp, ty)
    Lm = np.zeros((g.cell_total, g.cell_total)) if (want_l and rc0 == 0) else None
    ms = C.c_double(0.0)
    rc = product_lib().lib.sphg_rf_generate(C.byref(f), xp.shape[0], _dptr(xp), _iptr(ty), nreal, _dptr(zeta),
                                            fill, _dptr(value), _dptr(Lm) if Lm is not None else None,
                                            C.byref(ms))
    return int(rc), value, Lm, ms.value


def rf_generate_pair(f_fai: RfField, f_c: RfField, coe: float, xp: np.ndarray, ty: np.ndarray, zeta: np.ndarray,
                     fill: float = 0.0):
    """sphg_rf_generate_pair: (rc, value_fai[nreal][n], value_c[nreal][n]); zeta is
    (nreal, cell_total, 2)."""
    xp = np.ascontiguousarray(xp, dtype=np.float64)
    ty = np.ascontiguousarray(ty, np.int32)
    zeta = np.ascontiguousarray(zeta, dtype=np.float64)
    if zeta.ndim == 2:
        zeta = zeta[None]
    nreal = zeta.shape[0]
    va = np.zeros((nreal, xp.shape[0]))
    vc = np.zeros((nreal, xp.shape[0]))
    rc = product_lib().lib.sphg_rf_generate_pair(C.byref(f_fai), C.byref(f_c), coe, xp.shape[0], _dptr(xp),
                                                 _iptr(ty), nreal, _dptr(zeta), fill, _dptr(va), _dptr(vc))
    return int(rc), va, vc


def kl_field(deck: Deck, dist: int, mean: float, sd: float, rl, field: int = 1, coe_depth: float = 0.0) -> KLField:
    """K-L field for `deck` over the reference's domain (sphg_kl_domain) with its modes
    (sphg_kl_modes); the depth-trend references are the deck's soil / water maxima on the
    vertical axis the reference reads (y in 2D, the z entry in 3D)."""
    L = product_lib().lib
    f = KLField()
    f.ndim, f.dist, f.field = deck.ndim, dist, field
    f.mean_value, f.sd_value, f.coe_depth = mean, sd, coe_depth
    for k in range(3):
        f.rl[k] = float(rl[k]) if k < len(rl) else 0.0
    xp = np.ascontiguousarray(deck.xp, dtype=np.float64)
    ty = np.ascontiguousarray(deck.type, np.int32)
    rc = L.sphg_kl_domain(deck.ndim, deck.n, _dptr(xp), _iptr(ty), f.a, f.x_trans)
    if rc:
        raise RuntimeError(f"sphg_kl_domain: error {rc}")
    _, _, ph = deck.domain()
    ax = 1 if deck.ndim == 2 else 2
    f.soil_max = float(ph["soil"][1][ax])
    f.water_max = float(ph["water"][1][ax])
    rc = L.sphg_kl_modes(C.byref(f))
    if rc:
        raise RuntimeError(f"sphg_kl_modes: error {rc}")
    return f


class Rain(C.Structure):
    """sphg_rain: Para_Rain as the RP section gives it."""
    _fields_ = [("rtype", _I), ("nrain", _I), ("nhor", _I), ("press_factor", _D), ("drop_velocity", _D)]


class RVFlags(C.Structure):
    """sphg_rv_flags: the Flag_SpatialVariability members random_variable_generate reads."""
    _fields_ = [("flag_fai", _I), ("flag_c", _I), ("flag_faiw", _I), ("flag_cw", _I), ("coe_frict_cohesion", _D)]


class RVParam(C.Structure):
    """sphg_rv_param: distribution type (1 normal, 2 lognormal), mean and standard deviation."""
    _fields_ = [("type", _I), ("mean_value", _D), ("sd_value", _D)]


def rv_args(flags, coe, psv):
    """flags (fai, c, faiw, cw), coefficient and six (type, mean, sd) tuples as the C structs."""
    fl = RVFlags(*[int(v) for v in flags], float(coe))
    ps = (RVParam * 6)(*[RVParam(int(t), float(m), float(sd)) for (t, m, sd) in psv])
    return fl, ps


def random_variable_values(flags, coe, psv, zeta):
    """sphg_random_variable_values: the six homogeneous values (NaN where not flagged)."""
    fl, ps = rv_args(flags, coe, psv)
    out = (_D * 6)()
    rc = product_lib().lib.sphg_random_variable_values(C.byref(fl), ps, (_D * 2)(*zeta), out)
    if rc:
        raise RuntimeError(f"sphg_random_variable_values: error {rc}")
    return np.array(list(out))


def kl_corr_normals(fa: KLField, fb: KLField, coe: float, u: np.ndarray):
    """sphg_kl_corr_normals: the mixed standard normals (xi_a, xi_b) and the coefficient."""
    u = np.ascontiguousarray(u, dtype=np.float64).reshape(200, 2)
    xa, xb, cf = np.zeros(200), np.zeros(200), C.c_double()
    rc = product_lib().lib.sphg_kl_corr_normals(C.byref(fa), C.byref(fb), coe, _dptr(u), _dptr(xa), _dptr(xb),
                                                C.byref(cf))
    if rc:
        raise RuntimeError(f"sphg_kl_corr_normals: error {rc}")
    return xa, xb, cf.value


def kl_synthesize_batch(f: KLField, xp: np.ndarray, ty: np.ndarray, xi: np.ndarray, fill: float = 0.0):
    """Field values [nreal][n] of a realisation batch in the common domain (FP32 matrix-core GEMM)
    and the GEMM kernel time in ms."""
    L = product_lib().lib
    xp = np.ascontiguousarray(xp, dtype=np.float64)
    ty = np.ascontiguousarray(ty, np.int32)
    xi = np.ascontiguousarray(xi, dtype=np.float64)
    nreal = xi.shape[0]
    out = np.zeros((nreal, xp.shape[0]))
    ms = C.c_double(0.0)
    rc = L.sphg_kl_synthesize_batch(C.byref(f), xp.shape[0], _dptr(xp), _iptr(ty), nreal, _dptr(xi), fill,
                                    _dptr(out), C.byref(ms))
    if rc:
        raise RuntimeError(f"sphg_kl_synthesize_batch: error {rc}")
    return out, ms.value


def problem_from_deck(d: Deck) -> Problem:
    p = Problem()
    for k in "fla fls flr fli flv fleob flts flspv flknl ntd flbndy ndim fop inip loop step_regu dttype".split():
        setattr(p, k, int(getattr(d, k)))
    p.ntotal = d.n
    p.dr = d.dr
    p.dt = d.dt
    p.tt = d.dt * (d.loop + 1)  # CPP_2-initialf.cpp:277
    for i in range(3):
        p.xsph[i] = d.xsph[i]
        for j in range(2):
            p.art_vis[i][j] = d.art_vis[i][j]
    p.art_stre_coe = d.art_stre_coe
    p.threshold_min, p.threshold_max = d.threshold_min, d.threshold_max
    p.gx, p.gy, p.gz = d.gravity
    p.if_cal, p.if_move, p.coe_bndy = d.if_cal, d.if_move, d.coe_bndy
    for i in range(3):
        p.move_vx[i] = d.move_vx[i]
    p.interc, p.permtype, p.pers, p.perl, p.perm = d.interc, d.permtype, d.pers, d.perl, d.perm
    p.ss_type, p.ss_myu, p.ss_zeta, p.ss_acc_type = d.ss_type, d.ss_myu, d.ss_zeta, d.ss_acc_type
    for name, v in zip("flagem flagsb flagdp flagum flagam flagbui flagpl flagve flagfocc flagep".split(),
                       d.model_flags):
        setattr(p, name, int(v) if d.fls == 1 else 0)
    p.alphats = d.alphats if (d.fls == 1 and d.flts >= 1) else 0.0  # read with damp_coe (MS section)
    bmin, bmax, ph = d.domain()
    p.xmin, p.ymin, p.zmin = bmin
    p.xmax, p.ymax, p.zmax = bmax
    for i in range(3):
        p.soil_min[i], p.soil_max[i] = ph["soil"][0][i], ph["soil"][1][i]
        p.water_min[i], p.water_max[i] = ph["water"][0][i], ph["water"][1][i]
    return p


def materials_from_deck(d: Deck) -> Materials:
    m = Materials()
    for i, s in enumerate(d.soil_table()):
        for k in SOIL_FIELDS:
            setattr(m.soil[i], k, s[k])
    water, air = d.fluid_tables()
    for dst, src in list(zip(m.water, water)) + [(m.air, air)]:
        for k in FLUID_FIELDS:
            setattr(dst, k, int(src[k]) if k in ("bhtype", "press_type", "acc_type") else float(src[k]))
    s = d.structure_table()
    for k in STRUCT_FIELDS:
        setattr(m.structure, k, s[k])
    return m


# Particle member layout used by both the product and the oracle (ncomp per field)
PARTICLE_FIELDS = {
    "x": 3, "xp": 3, "vx": 3, "vxp": 3, "ax": 3, "axp": 3, "ux": 3, "strep": 3, "divx": 9, "interf": 3,
    "interfss": 3, "rho": 1, "rhop": 1, "mass": 1, "hl": 1, "pre": 1, "prep": 1, "e": 1, "eps": 6,
    "deps": 6, "sig": 6, "dsig": 6, "weps": 9, "adps": 6, "satu": 1, "dsatu": 1, "suct": 1, "cop": 1,
    "vsig": 1, "veps": 1, "divq": 1, "divr": 1, "meps": 1, "roue": 1, "rous": 1, "prew": 1, "prea": 1,
    "porosity": 1, "cd": 1, "poro_cop": 1, "beta": 9, "beta_ts": 9, "Rij": 9,
}
INT_FIELDS = ("type", "matype", "permtype", "etype")
# fields every step reads or writes (what must be uploaded for a restart)
STATE_FIELDS = ["xp", "vxp", "vx", "ax", "axp", "ux", "interf", "interfss", "rho", "rhop", "mass", "hl",
                "pre", "prep", "e", "eps", "sig", "adps", "satu", "dsatu", "suct", "cop", "vsig", "veps",
                "divq", "divr", "meps", "roue", "prew", "prea", "porosity", "cd", "poro_cop", "x"]


def _dptr(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _iptr(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_int))


class _Lib:
    """Binds one shared library exporting the context API under a prefix (sphg_ or or_)."""

    def __init__(self, path: str, prefix: str):
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} is not built")
        self.lib = C.CDLL(path)
        self.prefix = prefix
        L = self.lib
        for nm in ("set_field", "get_field"):
            f = getattr(L, prefix + nm)
            f.argtypes = [C.c_void_p, C.c_char_p, C.POINTER(C.c_double), _I, _I]
            f.restype = _I
        for nm in ("set_field_i", "get_field_i"):
            f = getattr(L, prefix + nm)
            f.argtypes = [C.c_void_p, C.c_char_p, C.POINTER(C.c_int), _I]
            f.restype = _I
        f = getattr(L, prefix + "get_neighbors")
        f.argtypes = [C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_int), _I]
        f.restype = _I
        f = getattr(L, prefix + "step")
        f.argtypes = [C.c_void_p, _I]
        f.restype = _I

    def fn(self, name):
        return getattr(self.lib, self.prefix + name)


class Sim:
    """A context of either library with the same field-level interface."""

    def __init__(self, lib: _Lib, ctx, n: int):
        self.L = lib
        self.ctx = ctx
        self.n = n

    def set(self, name: str, arr) -> None:
        if name in INT_FIELDS:
            a = np.ascontiguousarray(arr, dtype=np.int32)
            rc = self.L.fn("set_field_i")(self.ctx, name.encode(), _iptr(a), self.n)
        else:
            a = np.ascontiguousarray(arr, dtype=np.float64)
            nc = 1 if a.ndim == 1 else a.shape[1]
            rc = self.L.fn("set_field")(self.ctx, name.encode(), _dptr(a), self.n, nc)
        if rc:
            raise RuntimeError(f"set_field {name}: error {rc}")

    def get(self, name: str, ncomp: Optional[int] = None) -> np.ndarray:
        if name in INT_FIELDS:
            a = np.zeros(self.n, np.int32)
            rc = self.L.fn("get_field_i")(self.ctx, name.encode(), _iptr(a), self.n)
        else:
            nc = ncomp or PARTICLE_FIELDS.get(name, 3 if name.startswith("vx_") or name == "vel_bndy" else 1)
            a = np.zeros((self.n, nc) if nc > 1 else self.n, np.float64)
            rc = self.L.fn("get_field")(self.ctx, name.encode(), _dptr(a), self.n, nc)
        if rc:
            raise RuntimeError(f"get_field {name}: error {rc}")
        return a

    def load_state(self, state: Dict[str, np.ndarray], names=None) -> None:
        for k in INT_FIELDS:
            self.set(k, state[k])
        for k in (names or list(PARTICLE_FIELDS)):
            if k in state:
                self.set(k, state[k])
        for k, v in state.items():
            if k in ("ps_fai", "ps_c", "ps_cop", "ps_ds", "pf_fai", "pf_c"):
                self.set(k, v)

    def step(self, nsteps: int = 1) -> int:
        return int(self.L.fn("step")(self.ctx, nsteps))

    def neighbors(self):
        cnt = np.zeros(self.n, np.int32)
        ids = np.zeros((self.n, MAX_NB), np.int32)
        rc = self.L.fn("get_neighbors")(self.ctx, _iptr(cnt), _iptr(ids), self.n)
        if rc:
            raise RuntimeError(f"get_neighbors: error {rc}")
        return cnt, ids


# ---------------------------------------------------------------------- the product library
_product: Optional[_Lib] = None


def product_lib() -> _Lib:
    """Load the HIP library built in-tree; raises if it is missing (no fallback path exists)."""
    global _product
    if _product is None:
        lib = _Lib(LIB_PATH, "sphg_")
        L = lib.lib
        L.sphg_create.argtypes = [C.POINTER(C.c_void_p), C.POINTER(Problem), C.POINTER(Materials), _I, _I]
        L.sphg_create.restype = _I
        L.sphg_destroy.argtypes = [C.c_void_p]
        L.sphg_upload_done.argtypes = [C.c_void_p, _I]
        L.sphg_upload_done.restype = _I
        L.sphg_sync.argtypes = [C.c_void_p]
        L.sphg_sync.restype = _I
        L.sphg_set_step.argtypes = [C.c_void_p, _I]
        L.sphg_set_step.restype = _I
        L.sphg_set_time.argtypes = [C.c_void_p, _D]
        L.sphg_set_time.restype = _I
        L.sphg_get_time.argtypes = [C.c_void_p, C.POINTER(_D)]
        L.sphg_get_time.restype = _I
        L.sphg_set_dt_mode.argtypes = [C.c_void_p, _I, _D, _D, _D]
        L.sphg_set_dt_mode.restype = _I
        L.sphg_get_dt.argtypes = [C.c_void_p, C.POINTER(_D)]
        L.sphg_get_dt.restype = _I
        L.sphg_set_vibration.argtypes = [C.c_void_p, _D, _I, _D, _D, C.POINTER(_D)]
        L.sphg_set_vibration.restype = _I
        L.sphg_set_rain.argtypes = [C.c_void_p, C.POINTER(Rain)]
        L.sphg_set_rain.restype = _I
        L.sphg_set_excavation.argtypes = [C.c_void_p, _I, C.POINTER(_I)]
        L.sphg_set_excavation.restype = _I
        L.sphg_error_string.restype = C.c_char_p
        L.sphg_build_info.restype = C.c_char_p
        L.sphg_enable_timing.argtypes = [C.c_void_p, _I]
        L.sphg_kernel_times.argtypes = [C.c_void_p, C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_int), _I]
        L.sphg_kernel_times.restype = _I
        L.sphg_device_bytes.argtypes = [C.c_void_p, C.POINTER(C.c_size_t)]
        L.sphg_setting_par.argtypes = [C.POINTER(Problem), C.POINTER(Materials), _I] + \
            [C.POINTER(C.c_double), C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int)] + \
            [C.POINTER(C.c_double)] * 14
        L.sphg_setting_par.restype = _I
        L.sphg_kl_domain.argtypes = [_I, _I, C.POINTER(C.c_double), C.POINTER(C.c_int), _D * 3, _D * 3]
        L.sphg_kl_domain.restype = _I
        L.sphg_kl_modes.argtypes = [C.POINTER(KLField)]
        L.sphg_kl_modes.restype = _I
        L.sphg_kl_synthesize.argtypes = [C.c_void_p, C.POINTER(KLField), C.POINTER(C.c_double)]
        L.sphg_kl_synthesize.restype = _I
        L.sphg_kl_synthesize_batch.argtypes = [C.POINTER(KLField), _I, C.POINTER(C.c_double), C.POINTER(C.c_int),
                                               _I, C.POINTER(C.c_double), _D, C.POINTER(C.c_double),
                                               C.POINTER(C.c_double)]
        L.sphg_kl_synthesize_batch.restype = _I
        L.sphg_kl_synthesize_part.argtypes = [C.c_void_p, C.POINTER(KLField), C.POINTER(C.c_double), _I, _I]
        L.sphg_kl_synthesize_part.restype = _I
        L.sphg_kl_corr_normals.argtypes = [C.POINTER(KLField), C.POINTER(KLField), _D, C.POINTER(C.c_double),
                                           C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.sphg_kl_corr_normals.restype = _I
        L.sphg_kl_synthesize_corr_part.argtypes = [C.c_void_p, C.POINTER(KLField), C.POINTER(KLField), _D,
                                                   C.POINTER(C.c_double), _I, _I]
        L.sphg_kl_synthesize_corr_part.restype = _I
        L.sphg_random_variable_values.argtypes = [C.POINTER(RVFlags), C.POINTER(RVParam), _D * 2, _D * 6]
        L.sphg_random_variable_values.restype = _I
        L.sphg_random_variable_generate.argtypes = [C.c_void_p, C.POINTER(RVFlags), C.POINTER(RVParam), _D * 2,
                                                    _I, _I]
        L.sphg_random_variable_generate.restype = _I
        L.sphg_create_batch.argtypes = [C.POINTER(C.c_void_p), C.POINTER(Problem), C.POINTER(Materials), _I,
                                        C.POINTER(C.c_int), C.POINTER(Box), _I]
        L.sphg_create_batch.restype = _I
        L.sphg_step_async.argtypes = [C.c_void_p, _I]
        L.sphg_step_async.restype = _I
        L.sphg_poll.argtypes = [C.c_void_p]
        L.sphg_poll.restype = _I
        L.sphg_run.argtypes = [C.c_void_p, _I, _I, C.c_double, C.POINTER(C.c_double), OUTPUT_FN, C.c_void_p]
        L.sphg_run.restype = _I
        L.sphg_autotune.argtypes = [C.c_void_p, _I, C.POINTER(C.c_int)]
        L.sphg_autotune.restype = _I
        L.sphg_set_batch_widths.argtypes = [C.c_void_p, C.POINTER(C.c_int)]
        L.sphg_set_search_stage.argtypes = [C.c_void_p, _I]
        L.sphg_set_search_stage.restype = _I
        L.sphg_get_search_stage.argtypes = [C.c_void_p, C.POINTER(C.c_int)]
        L.sphg_get_search_stage.restype = _I
        L.sphg_set_list_order.argtypes = [C.c_void_p, _I]
        L.sphg_set_list_order.restype = _I
        L.sphg_set_batch_widths.restype = _I
        L.sphg_particle_count.argtypes = [C.c_void_p]
        L.sphg_particle_count.restype = _I
        L.sphg_write_odb_fields.argtypes = [C.c_char_p, _I, C.c_double, _I, C.POINTER(OdbFields)]
        L.sphg_write_odb_fields.restype = _I
        L.sphg_write_odb.argtypes = [C.c_void_p, C.c_char_p, _I, C.c_double, _I, _I]
        L.sphg_write_odb.restype = _I
        L.sphg_read_input.argtypes = [C.c_char_p, C.POINTER(Problem), C.POINTER(Materials), C.POINTER(C.c_int), _I,
                                      C.POINTER(C.c_double), C.POINTER(C.c_double)] + [C.POINTER(C.c_int)] * 4
        L.sphg_read_input.restype = _I
        L.sphg_rf_cells.argtypes = [C.POINTER(RfField), _I, C.POINTER(C.c_double), C.POINTER(C.c_int),
                                    C.POINTER(C.c_int)]
        L.sphg_rf_cells.restype = _I
        L.sphg_rf_generate.argtypes = [C.POINTER(RfField), _I, C.POINTER(C.c_double), C.POINTER(C.c_int), _I,
                                       C.POINTER(C.c_double), _D, C.POINTER(C.c_double), C.POINTER(C.c_double),
                                       C.POINTER(C.c_double)]
        L.sphg_rf_generate.restype = _I
        L.sphg_rf_generate_pair.argtypes = [C.POINTER(RfField), C.POINTER(RfField), _D, _I, C.POINTER(C.c_double),
                                            C.POINTER(C.c_int), _I, C.POINTER(C.c_double), _D,
                                            C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.sphg_rf_generate_pair.restype = _I
        for nm in ("save_state", "load_state"):
            f = getattr(L, "sphg_" + nm)
            f.argtypes = [C.c_void_p, C.c_char_p]
            f.restype = _I
        for nm in ("parttocell", "partisearching", "density", "strain", "strain_to_stress", "accel", "xsph",
                   "update"):
            f = getattr(L, "sphg_" + nm)
            f.argtypes = [C.c_void_p]
            f.restype = _I
        _product = lib
    return _product


OUTPUT_FN = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_int, C.c_double)

ODB_V3 = ("xp", "ux", "ax", "vx", "interf", "interfss", "strep")
ODB_V6 = ("sig", "eps", "adps")
ODB_S = ("pre", "prea", "prew", "vsig", "veps", "meps", "divq", "divr", "satu", "suct")


class OdbFields(C.Structure):
    """sphg_odb_fields: host arrays of one .odb table."""
    _fields_ = [(k, C.POINTER(C.c_double)) for k in ODB_V3 + ODB_V6 + ODB_S] + \
               [("type", C.POINTER(C.c_int)), ("matype", C.POINTER(C.c_int))]


def write_odb_fields(dirname: str, l: int, t: float, state: Dict[str, np.ndarray]) -> str:
    """sphg_write_odb_fields on host arrays (a state dict as the golden fixtures hold it); returns
    the path written."""
    keep = {}
    f = OdbFields()
    n = None
    for k, nc in [(k, 3) for k in ODB_V3] + [(k, 6) for k in ODB_V6] + [(k, 1) for k in ODB_S]:
        a = np.ascontiguousarray(np.asarray(state[k], dtype=np.float64).reshape(-1, nc))
        n = a.shape[0] if n is None else n
        if a.shape[0] != n:
            raise ValueError(f"{k}: {a.shape[0]} particles, expected {n}")
        keep[k] = a
        setattr(f, k, _dptr(a))
    for k in ("type", "matype"):
        a = np.ascontiguousarray(np.asarray(state[k], dtype=np.int32).reshape(-1))
        keep[k] = a
        setattr(f, k, _iptr(a))
    rc = product_lib().lib.sphg_write_odb_fields(dirname.encode(), l, t, n, C.byref(f))
    if rc:
        raise RuntimeError(f"sphg_write_odb_fields: error {rc}")
    return odb_path(dirname, l)


def read_input(path: str):
    """sphg_read_input: (Problem, Materials, particles) of an input.dat deck; particles holds xp, vxp
    (n x 3) and type, matype, permtype, etype (n)."""
    L = product_lib().lib
    pro, mat, n = Problem(), Materials(), C.c_int(0)
    rc = L.sphg_read_input(path.encode(), C.byref(pro), C.byref(mat), C.byref(n), 0, None, None, None, None,
                           None, None)
    if rc:
        raise RuntimeError(f"sphg_read_input: error {rc}")
    nn = n.value
    parts = {"xp": np.zeros((nn, 3)), "vxp": np.zeros((nn, 3))}
    for k in ("type", "matype", "permtype", "etype"):
        parts[k] = np.zeros(nn, np.int32)
    rc = L.sphg_read_input(path.encode(), C.byref(pro), C.byref(mat), C.byref(n), nn, _dptr(parts["xp"]),
                           _dptr(parts["vxp"]), *[_iptr(parts[k]) for k in ("type", "matype", "permtype", "etype")])
    if rc:
        raise RuntimeError(f"sphg_read_input: error {rc}")
    return pro, mat, parts


def odb_path(dirname: str, l: int) -> str:
    """File name output_total gives the table of step l (CPP_14-2-output.cpp:27-76)."""
    return os.path.join(dirname, f"data-{l:010d}.odb" if l < 100000000 else "data-.odb")


class GpuSim(Sim):
    """One simulation resident on the GPU (sphg_create ... sphg_destroy)."""

    def __init__(self, deck: Deck, precision: int = 0, capacity: Optional[int] = None):
        lib = product_lib()
        self._pro = problem_from_deck(deck)
        self._mat = materials_from_deck(deck)
        ctx = C.c_void_p()
        rc = lib.lib.sphg_create(C.byref(ctx), C.byref(self._pro), C.byref(self._mat), capacity or deck.n,
                                 precision)
        if rc:
            raise RuntimeError(f"sphg_create: error {rc} ({lib.lib.sphg_error_string(rc).decode()})")
        super().__init__(lib, ctx, deck.n)
        if deck.flv > 0:
            self.set_vibration(deck)
        self.set_dynamic_sets(deck)

    def set_dynamic_sets(self, deck: Deck) -> None:
        """sphg_set_rain / sphg_set_excavation from the deck's Para_Rain and Para_EC."""
        if deck.flr == 1:
            rc = self.L.lib.sphg_set_rain(self.ctx, C.byref(Rain(*deck.rain)))
            if rc:
                raise RuntimeError(f"sphg_set_rain: error {rc}")
        if deck.fleob == 1:
            ecp = np.ascontiguousarray(np.array(deck.ecp, dtype=np.int32).reshape(-1, 3))
            rc = self.L.lib.sphg_set_excavation(self.ctx, ecp.shape[0], _iptr(ecp))
            if rc:
                raise RuntimeError(f"sphg_set_excavation: error {rc}")

    def set_vibration(self, deck: Deck) -> None:
        """sphg_set_vibration from the deck's Para_Vibra fields and loads."""
        loads = np.ascontiguousarray(deck.vib_loads, dtype=np.float64)
        rc = self.L.lib.sphg_set_vibration(self.ctx, deck.vib_ttime, deck.vib_nsteps, deck.vib_sttime,
                                           deck.vib_edtime, _dptr(loads))
        if rc:
            raise RuntimeError(f"sphg_set_vibration: error {rc}")

    def set_dt_mode(self, mode: int, c_cour: float = 0.25, c_wave: float = 600.0, dt_max: float = 1.0) -> int:
        """sphg_set_dt_mode: 0 the constant Para_Pro::dt, 1 the device CFL min-reduction."""
        return int(self.L.lib.sphg_set_dt_mode(self.ctx, mode, c_cour, c_wave, dt_max))

    def dt(self) -> float:
        """sphg_get_dt: the time step of the last step."""
        v = C.c_double()
        rc = self.L.lib.sphg_get_dt(self.ctx, C.byref(v))
        if rc:
            raise RuntimeError(f"sphg_get_dt: error {rc}")
        return v.value

    def set_time(self, t: float) -> None:
        """sphg_set_time: the clock Para_Pro::t."""
        rc = self.L.lib.sphg_set_time(self.ctx, t)
        if rc:
            raise RuntimeError(f"sphg_set_time: error {rc}")

    def time(self) -> float:
        t = C.c_double()
        self.L.lib.sphg_get_time(self.ctx, C.byref(t))
        return t.value

    def upload_done(self):
        rc = self.L.lib.sphg_upload_done(self.ctx, self.n)
        if rc:
            raise RuntimeError(f"sphg_upload_done: error {rc}")

    def sync(self) -> int:
        return int(self.L.lib.sphg_sync(self.ctx))

    def run(self, loop: int, fop: int, tt: float, t: float = 0.0, output=None):
        """The reference time loop (sphg_run): output(step, t) after step 1, every fop-th step and at
        t >= tt; returns (error code, simulated time)."""
        cb = OUTPUT_FN(lambda user, l, tm: int(output(l, tm) or 0)) if output else OUTPUT_FN()
        tv = C.c_double(t)
        rc = self.L.lib.sphg_run(self.ctx, loop, fop, tt, C.byref(tv), cb, None)
        return int(rc), tv.value

    def set_step(self, l: int):
        self.L.lib.sphg_set_step(self.ctx, l)

    def write_odb(self, dirname: str, l: int, t: float, first: int = 0, count: Optional[int] = None) -> str:
        """sphg_write_odb: the reference's .odb table of particles first..first+count-1."""
        count = self.n - first if count is None else count
        rc = self.L.lib.sphg_write_odb(self.ctx, dirname.encode(), l, t, first, count)
        if rc:
            raise RuntimeError(f"sphg_write_odb: error {rc}")
        return odb_path(dirname, l)

    def autotune(self, steps: int = 5):
        """sphg_autotune: returns (error code, widths of density, normalisation, strain, XSPH)."""
        w = (C.c_int * 4)()
        rc = self.L.lib.sphg_autotune(self.ctx, steps, w)
        return int(rc), list(w)

    def set_list_order(self, reversed_: bool) -> int:
        """sphg_set_list_order: every neighbour sum in reversed list order (noise-floor runs)."""
        return int(self.L.lib.sphg_set_list_order(self.ctx, 1 if reversed_ else 0))

    def set_batch_widths(self, widths) -> int:
        """sphg_set_batch_widths (density, normalisation, strain, XSPH)."""
        return int(self.L.lib.sphg_set_batch_widths(self.ctx, (C.c_int * 4)(*widths)))

    def set_search_stage(self, stage: int) -> int:
        """sphg_set_search_stage: list entries the search stages in LDS (16 or 32)."""
        return int(self.L.lib.sphg_set_search_stage(self.ctx, stage))

    def search_stage(self) -> int:
        """sphg_get_search_stage."""
        v = C.c_int()
        rc = self.L.lib.sphg_get_search_stage(self.ctx, C.byref(v))
        if rc:
            raise RuntimeError(f"sphg_get_search_stage: error {rc}")
        return v.value

    def save_state(self, path: str) -> int:
        """sphg_save_state: checkpoint of the device state (returns the error code)."""
        return int(self.L.lib.sphg_save_state(self.ctx, path.encode()))

    def restore_state(self, path: str) -> int:
        """sphg_load_state into this (created and uploaded) context; returns the error code."""
        return int(self.L.lib.sphg_load_state(self.ctx, path.encode()))

    def call(self, name: str) -> int:
        return int(getattr(self.L.lib, "sphg_" + name)(self.ctx))

    def enable_timing(self, on=True):
        self.L.lib.sphg_enable_timing(self.ctx, 1 if on else 0)

    def kernel_times(self):
        cap = 64
        names = (C.c_char * 32 * cap)()
        ms = (C.c_double * cap)()
        cnt = (C.c_int * cap)()
        k = self.L.lib.sphg_kernel_times(self.ctx, names, ms, cnt, cap)
        return [(bytes(names[i]).split(b"\0", 1)[0].decode(), ms[i], cnt[i]) for i in range(k)]

    def kl_synthesize(self, f: "KLField", xi: np.ndarray) -> None:
        """sphg_kl_synthesize: K-L field into the per-particle property f.field."""
        xi = np.ascontiguousarray(xi, dtype=np.float64)
        rc = self.L.lib.sphg_kl_synthesize(self.ctx, C.byref(f), _dptr(xi))
        if rc:
            raise RuntimeError(f"sphg_kl_synthesize: error {rc}")

    def _range(self, r):
        if r is None:
            return 0, self.n
        a, b = int(self.offsets[r]), int(self.offsets[r + 1])
        return a, b - a

    def kl_synthesize_corr(self, fa: "KLField", fb: "KLField", coe: float, u: np.ndarray, r=None) -> int:
        """sphg_kl_synthesize_corr(_part): correlated fai / c pair from u[200][2]; returns the code."""
        u = np.ascontiguousarray(u, dtype=np.float64).reshape(200, 2)
        a, cnt = self._range(r)
        return self.L.lib.sphg_kl_synthesize_corr_part(self.ctx, C.byref(fa), C.byref(fb), coe, _dptr(u), a, cnt)

    def random_variable_generate(self, flags, coe, psv, zeta, r=None) -> int:
        """sphg_random_variable_generate over the context (or realisation r); returns the code."""
        fl, ps = rv_args(flags, coe, psv)
        a, cnt = self._range(r)
        return self.L.lib.sphg_random_variable_generate(self.ctx, C.byref(fl), ps, (_D * 2)(*zeta), a, cnt)

    def device_bytes(self) -> int:
        b = C.c_size_t()
        self.L.lib.sphg_device_bytes(self.ctx, C.byref(b))
        return int(b.value)

    def close(self):
        if self.ctx:
            self.L.lib.sphg_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class GpuBatch(GpuSim):
    """A realisation batch in one context (sphg_create_batch): decks of one problem family, each
    with its own particles, domain box and cell grid; particles concatenated realisation-major."""

    def __init__(self, decks, precision: int = 0):
        lib = product_lib()
        self.decks = list(decks)
        self._pro = problem_from_deck(self.decks[0])
        self._mat = materials_from_deck(self.decks[0])
        nr = len(self.decks)
        counts = (C.c_int * nr)(*[d.n for d in self.decks])
        boxes = (Box * nr)()
        for r, d in enumerate(self.decks):
            bmin, bmax, _ = d.domain()
            boxes[r].xmin, boxes[r].xmax = bmin[0], bmax[0]
            boxes[r].ymin, boxes[r].ymax = bmin[1], bmax[1]
            boxes[r].zmin, boxes[r].zmax = bmin[2], bmax[2]
        ctx = C.c_void_p()
        rc = lib.lib.sphg_create_batch(C.byref(ctx), C.byref(self._pro), C.byref(self._mat), nr, counts, boxes,
                                       precision)
        if rc:
            raise RuntimeError(f"sphg_create_batch: error {rc}")
        self.offsets = np.concatenate([[0], np.cumsum([d.n for d in self.decks])]).astype(np.int64)
        Sim.__init__(self, lib, ctx, int(self.offsets[-1]))
        if self.decks[0].flv > 0:
            self.set_vibration(self.decks[0])

    def kl_synthesize_part(self, r: int, f: "KLField", xi: np.ndarray) -> None:
        xi = np.ascontiguousarray(xi, dtype=np.float64)
        a, b = int(self.offsets[r]), int(self.offsets[r + 1])
        rc = self.L.lib.sphg_kl_synthesize_part(self.ctx, C.byref(f), _dptr(xi), a, b - a)
        if rc:
            raise RuntimeError(f"sphg_kl_synthesize_part: error {rc}")

    def part(self, name: str, r: int) -> np.ndarray:
        a, b = int(self.offsets[r]), int(self.offsets[r + 1])
        return self.get(name)[a:b]


def batch_state(decks) -> Dict[str, np.ndarray]:
    """Concatenated initial states (setting_par of every deck), realisation-major."""
    parts = [setting_par(d) for d in decks]
    return {k: np.concatenate([p[k] for p in parts]) for k in parts[0]}


def setting_par(deck: Deck) -> Dict[str, np.ndarray]:
    """Initial particle state via the library's host restatement of setting_vis_cons +
    setting_par (CPP_2-initialf.cpp:1021-1380). Host only; no GPU needed."""
    L = product_lib().lib
    pro = problem_from_deck(deck)
    mat = materials_from_deck(deck)
    n = deck.n
    xp = np.ascontiguousarray(deck.xp, dtype=np.float64)
    ty = np.ascontiguousarray(deck.type, np.int32)
    mt = np.ascontiguousarray(deck.matype, np.int32)
    et = np.ascontiguousarray(deck.etype, np.int32)
    out = {k: np.zeros(n) for k in ("rho", "rhop", "mass", "hl", "satu", "e", "porosity", "cd", "roue", "pre",
                                     "prep", "suct")}
    out["axp"] = np.zeros((n, 3))
    out["sig"] = np.zeros((n, 6))
    rc = L.sphg_setting_par(C.byref(pro), C.byref(mat), n, _dptr(xp), _iptr(ty), _iptr(mt), _iptr(et),
                            _dptr(out["rho"]), _dptr(out["rhop"]), _dptr(out["mass"]), _dptr(out["hl"]),
                            _dptr(out["satu"]), _dptr(out["e"]), _dptr(out["porosity"]), _dptr(out["axp"]),
                            _dptr(out["cd"]), _dptr(out["roue"]), _dptr(out["sig"]), _dptr(out["pre"]),
                            _dptr(out["prep"]), _dptr(out["suct"]))
    if rc:
        raise RuntimeError(f"sphg_setting_par: error {rc}")
    out["xp"] = xp.copy()
    out["vxp"] = np.ascontiguousarray(deck.vxp, dtype=np.float64).copy()
    out["type"], out["matype"], out["etype"] = ty.copy(), mt.copy(), et.copy()
    out["permtype"] = np.ascontiguousarray(deck.permtype_p, np.int32).copy()
    return out
