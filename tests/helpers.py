"""Test-side helpers: the oracle library binding and the reference link harness runner.

Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg use these (the oracle is
the checker, never the product)."""
from __future__ import annotations

import ctypes as C
import os
import shutil
import subprocess
import sys
import tempfile
from typing import Dict, Iterable, Optional

import numpy as np

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(REPO, "multiphase-sph-solver_amd")
if PKG not in sys.path:
    sys.path.insert(0, PKG)

from sphgpy.capi import (Problem, Materials, Sim, _Lib, problem_from_deck, materials_from_deck,  # noqa: E402
                         PARTICLE_FIELDS, INT_FIELDS)
from sphgpy.dumpio import read_dump  # noqa: E402

ORACLE_SO = os.path.join(REPO, "oracle", "_build", "libsph_oracle.so")
REF_HARNESS = os.path.join(REPO, "oracle", "_ref", "ref_harness")
GOLDEN = os.path.join(REPO, "tests", "golden")

_or: Optional[_Lib] = None


def oracle_lib() -> _Lib:
    global _or
    if _or is None:
        if not os.path.exists(ORACLE_SO):
            subprocess.run(["make", "-C", os.path.join(REPO, "oracle"), "oracle"], check=True,
                           capture_output=True)
        lib = _Lib(ORACLE_SO, "or_")
        L = lib.lib
        L.or_create.argtypes = [C.POINTER(Problem), C.POINTER(Materials), C.c_int]
        L.or_create.restype = C.c_void_p
        L.or_destroy.argtypes = [C.c_void_p]
        L.or_upload_done.argtypes = [C.c_void_p]
        L.or_setting_vis_cons.argtypes = [C.c_void_p]
        L.or_set_mode.argtypes = [C.c_void_p, C.c_int, C.c_int]
        L.or_set_step.argtypes = [C.c_void_p, C.c_int]
        L.or_set_time.argtypes = [C.c_void_p, C.c_double]
        L.or_set_vibration.argtypes = [C.c_void_p, C.c_double, C.c_int, C.c_double, C.c_double,
                                       C.POINTER(C.c_double)]
        L.or_set_vibration.restype = C.c_int
        L.or_set_rain.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double]
        L.or_set_rain.restype = C.c_int
        L.or_set_excavation.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_int)]
        L.or_set_excavation.restype = C.c_int
        L.or_phase.argtypes = [C.c_void_p, C.c_int]
        L.or_phase.restype = C.c_int
        L.or_pair_count.argtypes = [C.c_void_p]
        L.or_pair_count.restype = C.c_longlong
        _or = lib
    return _or


GS, JACOBI = 0, 1


class OracleSim(Sim):
    def __init__(self, deck, state: Dict[str, np.ndarray], mode: int = GS, nthreads: int = 1, step: int = 1):
        lib = oracle_lib()
        self._pro = problem_from_deck(deck)
        self._mat = materials_from_deck(deck)
        ctx = lib.lib.or_create(C.byref(self._pro), C.byref(self._mat), deck.n)
        super().__init__(lib, ctx, deck.n)
        if deck.flv > 0:
            loads = np.ascontiguousarray(deck.vib_loads, dtype=np.float64)
            lib.lib.or_set_vibration(ctx, deck.vib_ttime, deck.vib_nsteps, deck.vib_sttime, deck.vib_edtime,
                                     loads.ctypes.data_as(C.POINTER(C.c_double)))
        for k in INT_FIELDS:
            self.set(k, state[k])
        lib.lib.or_setting_vis_cons(ctx)
        self.load_state(state)
        for k, v in state.items():
            if k.startswith("ps_") or k.startswith("pf_") or k in ("vx_water", "vx_soil", "vx_air", "vx_struct", "vel_bndy"):
                self.set(k, v)
        if deck.flr == 1:
            rc = lib.lib.or_set_rain(ctx, *deck.rain)
            if rc:
                raise RuntimeError(f"or_set_rain: error {rc}")
        if deck.fleob == 1:
            ecp = np.ascontiguousarray(np.array(deck.ecp, dtype=np.int32).reshape(-1, 3))
            lib.lib.or_set_excavation(ctx, ecp.shape[0], ecp.ctypes.data_as(C.POINTER(C.c_int)))
        lib.lib.or_upload_done(ctx)
        lib.lib.or_set_mode(ctx, mode, nthreads)
        lib.lib.or_set_step(ctx, step)

    def phase(self, mask: int) -> int:
        return int(self.L.lib.or_phase(self.ctx, mask))

    def pair_count(self) -> int:
        return int(self.L.lib.or_pair_count(self.ctx))

    def close(self):
        if self.ctx:
            self.L.lib.or_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def have_ref() -> bool:
    return os.path.exists(REF_HARNESS)


def run_ref(deck, steps: int, fop: int, workdir: Optional[str] = None, ntd: int = 1) -> Dict[int, dict]:
    """Run the reference solver (link harness) on `deck` for `steps` steps, dumping every `fop`
    steps (and step 1). Returns {step: fields} including step 0 = state after setting_par."""
    d = deck.copy()
    d.loop, d.fop, d.ntd = steps, fop, ntd
    wd = workdir or tempfile.mkdtemp(prefix="sphref_")
    os.makedirs(wd, exist_ok=True)
    d.write_input_dat(os.path.join(wd, "input.dat"))
    r = subprocess.run([REF_HARNESS, wd], capture_output=True)
    if r.returncode != 0:
        raise RuntimeError(f"reference harness failed rc={r.returncode}: {r.stdout[-400:]!r}")
    out = {}
    for fn in sorted(os.listdir(wd)):
        if fn.startswith("dump_") and fn.endswith(".bin"):
            f, step, _ = read_dump(os.path.join(wd, fn))
            out[step] = f
    if workdir is None:
        shutil.rmtree(wd, ignore_errors=True)
    return out


def max_rel(a: np.ndarray, b: np.ndarray) -> float:
    """Field-max normalised difference max|a-b| / max|b| (0 when both are all zero)."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    den = np.max(np.abs(b))
    num = np.max(np.abs(a - b)) if a.size else 0.0
    if den == 0.0:
        return float(num)
    return float(num / den)


# ---------------------------------------------------------------------------------- K-L
KL_TERMS = 200


def oracle_kl(ndim, xp, ty, a, rl, x_trans, xi, dist, mean, sd, coe_depth, soil_max, water_max, field, value):
    """Oracle K-L (oracle/kl_oracle.c): returns wi[200][3], eigval[200], order[200][3], value[n], rc."""
    L = oracle_lib().lib
    D3 = C.c_double * 3
    wi = np.zeros((KL_TERMS, 3))
    ev = np.zeros(KL_TERMS)
    od = np.zeros((KL_TERMS, 3), np.int32)
    dp = lambda a_: a_.ctypes.data_as(C.POINTER(C.c_double))  # noqa: E731
    ip = lambda a_: a_.ctypes.data_as(C.POINTER(C.c_int))  # noqa: E731
    rcw = L.or_kl_solve_wi(dp(wi), D3(*a), D3(*rl), ndim)
    if rcw:
        return wi, ev, od, np.ascontiguousarray(value, dtype=np.float64).copy(), rcw
    L.or_kl_solve_ramdai(dp(ev), dp(wi), D3(*rl), ip(od), ndim)
    xp = np.ascontiguousarray(xp, dtype=np.float64)
    ty = np.ascontiguousarray(ty, np.int32)
    xi = np.ascontiguousarray(xi, dtype=np.float64)
    out = np.ascontiguousarray(value, dtype=np.float64).copy()
    L.or_kl_synthesize.restype = C.c_int
    L.or_kl_synthesize.argtypes = [C.c_int, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_int),
                                   C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(C.c_int),
                                   C.POINTER(C.c_double), D3, D3, C.c_int, C.c_double, C.c_double, C.c_double,
                                   C.c_double, C.c_double, C.c_int, C.POINTER(C.c_double)]
    rc = L.or_kl_synthesize(ndim, xp.shape[0], dp(xp), ip(ty), dp(wi), dp(ev), ip(od), dp(xi), D3(*x_trans),
                            D3(*a), dist, mean, sd, coe_depth, soil_max, water_max, field, dp(out))
    return wi, ev, od, out, rc


def run_ref_kl(ndim, xp, ty, a, rl, x_trans, xi, dist, mean, sd, coe_depth, soil_max, water_max, field, value):
    """The reference's own solve_wi / solve_ramdai / solve_h_2D|3D through the link harness."""
    import struct
    n = xp.shape[0]
    wd = tempfile.mkdtemp(prefix="sphkl_")
    try:
        with open(os.path.join(wd, "kl_in.bin"), "wb") as f:
            f.write(struct.pack("<ii", ndim, n))
            f.write(np.asarray(a, np.float64).tobytes() + np.asarray(rl, np.float64).tobytes() +
                    np.asarray(x_trans, np.float64).tobytes())
            f.write(struct.pack("<i", dist) + struct.pack("<5d", mean, sd, coe_depth, soil_max, water_max) +
                    struct.pack("<i", field))
            f.write(np.asarray(xi, np.float64).tobytes())
            f.write(np.ascontiguousarray(xp, np.float64).tobytes() + np.ascontiguousarray(ty, np.int32).tobytes() +
                    np.ascontiguousarray(value, np.float64).tobytes())
        r = subprocess.run([REF_HARNESS, "--kl", wd], capture_output=True)
        if r.returncode != 0:
            raise RuntimeError(f"reference harness --kl failed rc={r.returncode}")
        raw = open(os.path.join(wd, "kl_out.bin"), "rb").read()
        o = 0
        wi = np.frombuffer(raw, np.float64, 600, o).reshape(200, 3).copy(); o += 4800
        ev = np.frombuffer(raw, np.float64, 200, o).copy(); o += 1600
        od = np.frombuffer(raw, np.int32, 600, o).reshape(200, 3).copy(); o += 2400
        val = np.frombuffer(raw, np.float64, n, o).copy(); o += 8 * n
        rc = int(np.frombuffer(raw, np.int32, 1, o)[0])
        return wi, ev, od, val, rc
    finally:
        shutil.rmtree(wd, ignore_errors=True)


# ------------------------------------------------------------ correlation-matrix random fields
def oracle_rf(ndim, xp, ty, f, zeta, fill=0.0, want_l=False, cells_only=False):
    """Oracle random field (oracle/rf_oracle.c) for an RfField f (the product's parameter struct,
    read field by field): returns rc, value[n], cells dict, L (if want_l), E; with cells_only the
    cell partition alone (rc, None, cells, None, None)."""
    L = oracle_lib().lib
    dp = lambda a_: a_.ctypes.data_as(C.POINTER(C.c_double))  # noqa: E731
    ip = lambda a_: a_.ctypes.data_as(C.POINTER(C.c_int))  # noqa: E731
    xp = np.ascontiguousarray(xp, dtype=np.float64)
    ty = np.ascontiguousarray(ty, np.int32)
    n = xp.shape[0]
    nc = np.zeros(3, np.int32)
    base = np.zeros(3)
    cdx = C.c_double(0.0)
    tot = C.c_int(0)
    cell_of = np.zeros(n, np.int32)
    box = np.array(list(f.box_nb), dtype=np.float64)
    L.or_rf_cells.restype = C.c_int
    rc = L.or_rf_cells(ndim, n, dp(xp), ip(ty), C.c_double(f.cell_ndx), C.c_double(f.dr), dp(box), ip(nc),
                       dp(base), C.byref(cdx), C.byref(tot), ip(cell_of))
    cells = {"nc": nc, "base": base, "cell_dx": cdx.value, "total": tot.value, "cell_of": cell_of}
    if rc or cells_only:
        return rc, None, cells, None, None
    total = tot.value
    A = np.zeros((total, total))
    rl = np.array(list(f.rl), dtype=np.float64)
    L.or_rf_corr(ndim, f.acf, dp(rl), C.c_double(f.beta), C.c_double(f.beta_1), ip(nc), dp(base), cdx, total,
                 dp(A))
    Lm = np.zeros((total, total))
    L.or_rf_cholesky(dp(A), dp(Lm), total)
    zeta = np.ascontiguousarray(zeta, dtype=np.float64)
    E = np.zeros(total)
    value = np.zeros(n)
    D = C.c_double
    L.or_rf_values(ndim, f.dist, f.field, D(f.mean_value), D(f.sd_value), D(f.coe_depth), D(f.soil_max),
                   D(f.water_max), ip(nc), dp(base), cdx, total, dp(Lm), dp(zeta), n, ip(ty), ip(cell_of), D(fill),
                   dp(E), dp(value))
    return 0, value, cells, (Lm if want_l else None), E


def oracle_chol(A):
    L = oracle_lib().lib
    A = np.ascontiguousarray(A, dtype=np.float64)
    out = np.zeros_like(A)
    L.or_rf_cholesky(A.ctypes.data_as(C.POINTER(C.c_double)), out.ctypes.data_as(C.POINTER(C.c_double)),
                     A.shape[0])
    return out


def run_ref_chol(A):
    """The reference's own Cholesky (CPP_17-3-Nonstationary-RF.cpp:45-74) through the link harness,
    serial (one OpenMP thread: its diagonal sum is an OpenMP reduction)."""
    import struct
    wd = tempfile.mkdtemp(prefix="sphchol_")
    try:
        A = np.ascontiguousarray(A, dtype="<f8")
        with open(os.path.join(wd, "chol_in.bin"), "wb") as f:
            f.write(struct.pack("<i", A.shape[0]) + A.tobytes())
        env = dict(os.environ, OMP_NUM_THREADS="1")
        r = subprocess.run([REF_HARNESS, "--chol", wd], capture_output=True, env=env)
        if r.returncode != 0:
            raise RuntimeError(f"reference harness --chol failed rc={r.returncode}")
        return np.fromfile(os.path.join(wd, "chol_out.bin"), "<f8").reshape(A.shape)
    finally:
        shutil.rmtree(wd, ignore_errors=True)


def run_ref_rf(ndim, xp, ty, dr, box_nb, soil_max3, water_max3, flags, cell_ndx, flag_auto, beta, beta_1, psv,
               init, correlated=False, coe=0.0, kind=None, files=None):
    """The reference's own nonstationary_random_2D/3D (or, correlated, its
    nonstationary_correlated_random_2D/3D with coe_frict_cohesion = coe) through the link harness;
    kind "klc" runs SpatialVariables_Generate_2D/3D_corr, "kln" SpatialVariables_Generate_2D/3D and
    "rv" random_variable_generate and "rfi" random_field_input (files: {name: text} written into
    the working directory first, Parameters.txt for "rfi") on the same inputs instead. psv: six tuples (type, mean, sd,
    rl[3], coe_depth); init: six n-arrays (soil fai c cop ds, fluid fai c). Returns (six n-arrays
    after, rc). Reproducible only where the values do not depend on the clock-seeded normals."""
    if kind is None:
        kind = "rfc" if correlated else "rf"
    import struct
    wd = tempfile.mkdtemp(prefix="sphrf_")
    try:
        n = xp.shape[0]
        for fname, text in (files or {}).items():
            with open(os.path.join(wd, fname), "w") as f:
                f.write(text)
        with open(os.path.join(wd, "rf_in.bin"), "wb") as f:
            f.write(struct.pack("<ii", ndim, n) + struct.pack("<d", dr))
            f.write(np.asarray(box_nb, "<f8").tobytes() + np.asarray(soil_max3, "<f8").tobytes() +
                    np.asarray(water_max3, "<f8").tobytes())
            f.write(np.asarray(flags, "<i4").tobytes() + struct.pack("<d", cell_ndx) + struct.pack("<i", flag_auto) +
                    struct.pack("<dd", beta, beta_1))
            for (t, mean, sd, rl, cdep) in psv:
                f.write(struct.pack("<i", t) + struct.pack("<6d", mean, sd, rl[0], rl[1], rl[2], cdep))
            f.write(np.ascontiguousarray(xp, "<f8").tobytes() + np.ascontiguousarray(ty, "<i4").tobytes())
            for a in init:
                f.write(np.ascontiguousarray(a, "<f8").tobytes())
            if kind != "rf":
                f.write(struct.pack("<d", coe))
        env = dict(os.environ, OMP_NUM_THREADS="1")
        r = subprocess.run([REF_HARNESS, "--" + kind, wd], capture_output=True, env=env, cwd=wd)
        if r.returncode != 0:
            raise RuntimeError(f"reference harness --rf failed rc={r.returncode}")
        raw = open(os.path.join(wd, "rf_out.bin"), "rb").read()
        vals = np.frombuffer(raw, "<f8", 6 * n).reshape(6, n).copy()
        rc = int(np.frombuffer(raw, "<i4", 1, 6 * n * 8)[0])
        if kind in ("klc", "kln"):  # and the half-widths a of the K-L domain it set, per property
            return vals, rc, np.frombuffer(raw, "<f8", 18, 6 * n * 8 + 4).reshape(6, 3).copy()
        return vals, rc
    finally:
        shutil.rmtree(wd, ignore_errors=True)


def oracle_rf_pair(ndim, xp, ty, f_fai, f_c, coe, zeta1, fill=0.0):
    """Oracle correlated pair (or_rf_pair_values) with the geometry of f_fai: returns rc, value_fai,
    value_c. zeta1 is (cell_total, 2)."""
    L = oracle_lib().lib
    dp = lambda a_: a_.ctypes.data_as(C.POINTER(C.c_double))  # noqa: E731
    ip = lambda a_: a_.ctypes.data_as(C.POINTER(C.c_int))  # noqa: E731
    rc, _, cells, Lm, _ = oracle_rf(ndim, xp, ty, f_fai, np.zeros(15000), want_l=True)
    if rc:
        return rc, None, None
    xp = np.ascontiguousarray(xp, dtype=np.float64)
    ty = np.ascontiguousarray(ty, np.int32)
    n = xp.shape[0]
    phase = 0 if f_fai.field == 0 else 1
    D = C.c_double
    D2 = C.c_double * 2
    zeta1 = np.ascontiguousarray(zeta1, dtype=np.float64)
    va, vc = np.zeros(n), np.zeros(n)
    nc = np.ascontiguousarray(cells["nc"], np.int32)
    base = np.ascontiguousarray(cells["base"], np.float64)
    L.or_rf_pair_values(ndim, phase, f_fai.dist, D(coe), D2(f_fai.mean_value, f_c.mean_value),
                        D2(f_fai.sd_value, f_c.sd_value), D2(f_fai.coe_depth, f_c.coe_depth), D(f_fai.soil_max),
                        D(f_fai.water_max), ip(nc), dp(base), D(cells["cell_dx"]), cells["total"], dp(Lm), dp(zeta1),
                        n, ip(ty), ip(cells["cell_of"]), D(fill), dp(va), dp(vc))
    return 0, va, vc


def oracle_kl_corr_normals(dist_a, mean_a, sd_a, mean_b, sd_b, coe, u):
    """Oracle mixing of the flspv 2 pair (or_kl_corr_normals): returns xi_a, xi_b, coefficient."""
    L = oracle_lib().lib
    D = C.c_double
    u = np.ascontiguousarray(u, dtype=np.float64).reshape(200, 2)
    xa, xb, cf = np.zeros(200), np.zeros(200), D()
    dp = lambda a_: a_.ctypes.data_as(C.POINTER(C.c_double))  # noqa: E731
    L.or_kl_corr_normals(dist_a, D(mean_a), D(sd_a), D(mean_b), D(sd_b), D(coe), dp(u), dp(xa), dp(xb), C.byref(cf))
    return xa, xb, cf.value


def oracle_random_variable(flags, coe, psv, zeta, ty, init):
    """Oracle random_variable_generate (or_random_variable): flags (fai, c, faiw, cw), psv six
    (type, mean, sd), zeta[2]; init six n-arrays; returns the six n-arrays after."""
    L = oracle_lib().lib
    D = C.c_double
    ty = np.ascontiguousarray(ty, np.int32)
    n = ty.shape[0]
    vals = np.ascontiguousarray(np.array(init, dtype=np.float64).reshape(6, n)).copy()
    dist = (C.c_int * 6)(*[int(p[0]) for p in psv])
    mean = (D * 6)(*[float(p[1]) for p in psv])
    sd = (D * 6)(*[float(p[2]) for p in psv])
    L.or_random_variable((C.c_int * 4)(*[int(v) for v in flags]), D(coe), dist, mean, sd, (D * 2)(*zeta), n,
                         ty.ctypes.data_as(C.POINTER(C.c_int)), vals.ctypes.data_as(C.POINTER(C.c_double)))
    return vals
