"""The reference-side binding (include/sphg_reference.hpp and the snippets of INTEGRATION.md)
compiled against the reference's own headers, where the reference checkout is present.

- The header's static_asserts check that Particle (Class_particle.h:13-103), Para_Soil, Para_Fluid and
  Para_Structure have the C records' layout member for member (a mismatch fails the build).
- INTEGRATION.md's C++ snippets compile inside a function that declares the variables
  SPH_Calculation holds (CPP_1-SPHLoops.cpp:67-378), with the same names and types.
- make_problem / make_materials, linked with the reference's own constructors of its parameter
  classes, copy every value they are given (distinct values per member, checked one by one).
"""
import os
import re
import subprocess

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = "/root/reference"
OBJ = os.path.join(REPO, "oracle", "_ref", "obj")

pytestmark = pytest.mark.skipif(not os.path.exists(os.path.join(REF, "Class_particle.h")),
                                reason="reference headers not present")

HEAD = """
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include "Class_Functions.h"
#include "sphg_reference.hpp"
"""

DRIVER_ARGS = """Para_FL *pFlag, Para_Pro *pPPro, Para_GF *pPGf, Para_Boundary *pPBou, Para_Inter *pPInter,
                  Para_SSInter *pSSInter, Para_Model *pMatter, Cell_Con *pCellc, Para_Soil *pSoil, Para_Fluid *pWater,
                  Para_Fluid *pAir, Para_Structure *pStruct, Particle *pPar, Para_Soil *pParti_ConsPara, char *argv,
                  FILE *flog"""


def _snippets():
    text = open(os.path.join(REPO, "INTEGRATION.md")).read()
    return re.findall(r"```cpp\n(.*?)```", text, re.S)


def _compile(src, tmp_path, link=None):
    f = tmp_path / "t.cpp"
    f.write_text(src)
    cmd = ["g++", "-std=c++14", "-w", f"-I{REF}", f"-I{os.path.join(REPO, 'include')}"]
    if link is None:
        cmd += ["-fsyntax-only", str(f)]
    else:
        cmd += [str(f)] + link + ["-o", str(tmp_path / "t")]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]


def test_integration_snippets_compile_against_reference_headers(tmp_path):
    blocks = _snippets()
    assert len(blocks) == 3
    init, loop, output = blocks
    init = init.replace('#include "sphg_reference.hpp"', "")
    file_scope, _, in_body = output.partition("// ...")
    src = HEAD + file_scope + f"""
int gpu_path({DRIVER_ARGS}) {{
    double threshold_min = 0.1, threshold_max = 0.12;
    int err_t = 0;
    clErr_Fun ItErrD;
    clOutput_Fun ItOutput;
{init}
{loop}
{in_body}
    sphg_destroy(gpu);
    return err_t;
}}
"""
    _compile(src, tmp_path)


@pytest.mark.skipif(not os.path.exists(os.path.join(OBJ, "Class_Problem.o")), reason="reference objects not built")
def test_make_problem_and_materials_copy_every_member(tmp_path):
    src = HEAD + r"""
#define CHK(a, b) do { if (!((a) == (b))) { printf("mismatch %s\n", #a); return 1; } } while (0)
int main() {
    Para_FL fl; Para_Pro pp; Para_GF gf; Para_Boundary bo; Para_Inter in; Para_SSInter ss; Para_Model mo; Cell_Con cc;
    fl.fla = 1; fl.fls = 2; fl.flr = 3; fl.fli = 4; fl.flv = 5; fl.fleob = 6; fl.flts = 7; fl.flspv = 8;
    fl.flknl = 9; fl.ntd = 10; fl.flbndy = 11;
    pp.ndim = 3; pp.fop = 12; pp.inip = 13; pp.loop = 14; pp.ntotal = 15; pp.step_regu = 16; pp.dttype = 2;
    pp.dr = 0.5; pp.dt = 0.25; pp.tt = 9.75; pp.art_stre_coe = 0.375;
    for (int k = 0; k < 3; k++) { pp.xsph[k] = 1.5 + k; pp.art_vis[k][0] = 2.5 + k; pp.art_vis[k][1] = 3.5 + k;
                                  bo.move_vx[k] = 4.5 + k; cc.soil_min[k] = 5.5 + k; cc.soil_max[k] = 6.5 + k;
                                  cc.water_min[k] = 7.5 + k; cc.water_max[k] = 8.5 + k; }
    gf.gx = 0.125; gf.gy = -9.5; gf.gz = 0.0625;
    bo.if_cal = 17; bo.if_move = 18; bo.coe_bndy = 0.03125;
    in.interc = 19.5; in.permtype = 20; in.pers = 21.5; in.perl = 22.5; in.perm = 23.5;
    ss.type = 24; ss.myu = 25.5; ss.zeta = 26.5; ss.acc_type = 27;
    mo.flagem = 28; mo.flagsb = 29; mo.flagdp = 30; mo.flagum = 31; mo.flagam = 32; mo.flagbui = 33; mo.flagpl = 34;
    mo.flagve = 35; mo.flagfocc = 36; mo.flagep = 37; mo.alphats = 38.5;
    cc.xmin = -1.5; cc.xmax = 39.5; cc.ymin = -2.5; cc.ymax = 40.5; cc.zmin = -3.5; cc.zmax = 41.5;
    sphg_problem p = sphg_ref::make_problem(fl, pp, 0.0078125, 0.015625, gf, bo, in, ss, mo, cc);
    CHK(p.fla, 1); CHK(p.fls, 2); CHK(p.flr, 3); CHK(p.fli, 4); CHK(p.flv, 5); CHK(p.fleob, 6); CHK(p.flts, 7);
    CHK(p.flspv, 8); CHK(p.flknl, 9); CHK(p.ntd, 10); CHK(p.flbndy, 11);
    CHK(p.ndim, 3); CHK(p.fop, 12); CHK(p.inip, 13); CHK(p.loop, 14); CHK(p.ntotal, 15); CHK(p.step_regu, 16);
    CHK(p.dttype, 2); CHK(p.dr, 0.5); CHK(p.dt, 0.25); CHK(p.tt, 9.75); CHK(p.art_stre_coe, 0.375);
    for (int k = 0; k < 3; k++) { CHK(p.xsph[k], 1.5 + k); CHK(p.art_vis[k][0], 2.5 + k); CHK(p.art_vis[k][1], 3.5 + k);
                                  CHK(p.move_vx[k], 4.5 + k); CHK(p.soil_min[k], 5.5 + k); CHK(p.soil_max[k], 6.5 + k);
                                  CHK(p.water_min[k], 7.5 + k); CHK(p.water_max[k], 8.5 + k); }
    CHK(p.threshold_min, 0.0078125); CHK(p.threshold_max, 0.015625);
    CHK(p.gx, 0.125); CHK(p.gy, -9.5); CHK(p.gz, 0.0625);
    CHK(p.if_cal, 17); CHK(p.if_move, 18); CHK(p.coe_bndy, 0.03125);
    CHK(p.interc, 19.5); CHK(p.permtype, 20); CHK(p.pers, 21.5); CHK(p.perl, 22.5); CHK(p.perm, 23.5);
    CHK(p.ss_type, 24); CHK(p.ss_myu, 25.5); CHK(p.ss_zeta, 26.5); CHK(p.ss_acc_type, 27);
    CHK(p.flagem, 28); CHK(p.flagsb, 29); CHK(p.flagdp, 30); CHK(p.flagum, 31); CHK(p.flagam, 32);
    CHK(p.flagbui, 33); CHK(p.flagpl, 34); CHK(p.flagve, 35); CHK(p.flagfocc, 36); CHK(p.flagep, 37);
    CHK(p.alphats, 38.5);
    CHK(p.xmin, -1.5); CHK(p.xmax, 39.5); CHK(p.ymin, -2.5); CHK(p.ymax, 40.5); CHK(p.zmin, -3.5); CHK(p.zmax, 41.5);
    // materials: every double of every record distinct
    Para_Soil soil[30]; Para_Fluid water[3], air; Para_Structure st;
    double v = 1.0;
    for (int k = 0; k < 30; k++) { double *d = reinterpret_cast<double *>(&soil[k]);
        for (size_t q = 0; q < sizeof(Para_Soil) / 8; q++) d[q] = (v += 1.0); }
    for (int k = 0; k < 3; k++) { water[k].vis = (v += 1); water[k].c = (v += 1); water[k].acc_type = k + 1;
                                  water[k].strain_min = (v += 1); }
    air.paa0 = (v += 1); air.press_type = 5; st.e = (v += 1); st.cop = (v += 1);
    sphg_materials m = sphg_ref::make_materials(soil, water, air, st);
    for (int k = 0; k < 30; k++) {
        CHK(m.soil[k].e, soil[k].e); CHK(m.soil[k].poi, soil[k].poi); CHK(m.soil[k].c, soil[k].c);
        CHK(m.soil[k].fai, soil[k].fai); CHK(m.soil[k].sme0, soil[k].sme0); CHK(m.soil[k].ds, soil[k].ds);
        CHK(m.soil[k].damp_coe, soil[k].damp_coe); CHK(m.soil[k].vpu, soil[k].vpu); }
    for (int k = 0; k < 3; k++) { CHK(m.water[k].vis, water[k].vis); CHK(m.water[k].c, water[k].c);
                                  CHK(m.water[k].acc_type, k + 1); CHK(m.water[k].strain_min, water[k].strain_min); }
    CHK(m.air.paa0, air.paa0); CHK(m.air.press_type, 5); CHK(m.structure.e, st.e); CHK(m.structure.cop, st.cop);
    printf("binding ok\n");
    return 0;
}
"""
    objs = [os.path.join(OBJ, o) for o in ("Class_Problem.o", "Class_ParticleProperty.o", "Class_NNPS.o",
                                           "Class_particle.o")]
    _compile(src, tmp_path, link=objs + ["-fopenmp", "-lm"])
    r = subprocess.run([str(tmp_path / "t")], capture_output=True, text=True)
    assert r.returncode == 0 and "binding ok" in r.stdout, r.stdout + r.stderr
