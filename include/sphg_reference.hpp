/*
 * sphg_reference.hpp — the reference driver's side of the binding: conversions from the reference
 * solver's own parameter, material and particle classes to the C ABI of include/sphg.h. Meant to
 * be included by CPP_1-SPHLoops.cpp (SPH_Calculation) in the reference tree, after its own headers
 * (Class_Functions.h brings Class_particle.h, Class_Problem.h, Class_ParticleProperty.h and
 * Class_NNPS.h); it adds no code to the library itself.
 *
 * The layout checks below fail the build if the reference's Particle, Para_Soil, Para_Fluid or
 * Para_Structure ever stops matching the C mirror member for member.
 */
#ifndef SPHG_REFERENCE_HPP_INCLUDED
#define SPHG_REFERENCE_HPP_INCLUDED

#include <cstddef>

#include "sphg.h"

namespace sphg_ref {

#define SPHG_SAME(A, B, f) static_assert(offsetof(A, f) == offsetof(B, f), "member " #f " of " #A " moved")
static_assert(sizeof(Particle) == sizeof(sphg_particle), "Particle (Class_particle.h:13-103) changed size");
SPHG_SAME(Particle, sphg_particle, x); SPHG_SAME(Particle, sphg_particle, xp); SPHG_SAME(Particle, sphg_particle, vx);
SPHG_SAME(Particle, sphg_particle, vxp); SPHG_SAME(Particle, sphg_particle, ax); SPHG_SAME(Particle, sphg_particle, axp);
SPHG_SAME(Particle, sphg_particle, ux); SPHG_SAME(Particle, sphg_particle, strep);
SPHG_SAME(Particle, sphg_particle, divx); SPHG_SAME(Particle, sphg_particle, interf);
SPHG_SAME(Particle, sphg_particle, interfss); SPHG_SAME(Particle, sphg_particle, rho);
SPHG_SAME(Particle, sphg_particle, rhop); SPHG_SAME(Particle, sphg_particle, mass); SPHG_SAME(Particle, sphg_particle, hl);
SPHG_SAME(Particle, sphg_particle, pre); SPHG_SAME(Particle, sphg_particle, prep); SPHG_SAME(Particle, sphg_particle, e);
SPHG_SAME(Particle, sphg_particle, eps); SPHG_SAME(Particle, sphg_particle, deps); SPHG_SAME(Particle, sphg_particle, sig);
SPHG_SAME(Particle, sphg_particle, dsig); SPHG_SAME(Particle, sphg_particle, ff_sig);
SPHG_SAME(Particle, sphg_particle, weps); SPHG_SAME(Particle, sphg_particle, adps);
SPHG_SAME(Particle, sphg_particle, tssig); SPHG_SAME(Particle, sphg_particle, tsqc);
SPHG_SAME(Particle, sphg_particle, satu); SPHG_SAME(Particle, sphg_particle, dsatu);
SPHG_SAME(Particle, sphg_particle, suct); SPHG_SAME(Particle, sphg_particle, cop); SPHG_SAME(Particle, sphg_particle, vsig);
SPHG_SAME(Particle, sphg_particle, veps); SPHG_SAME(Particle, sphg_particle, divq); SPHG_SAME(Particle, sphg_particle, divr);
SPHG_SAME(Particle, sphg_particle, meps); SPHG_SAME(Particle, sphg_particle, roue); SPHG_SAME(Particle, sphg_particle, rous);
SPHG_SAME(Particle, sphg_particle, zeta); SPHG_SAME(Particle, sphg_particle, rsta); SPHG_SAME(Particle, sphg_particle, prew);
SPHG_SAME(Particle, sphg_particle, prea); SPHG_SAME(Particle, sphg_particle, porosity); SPHG_SAME(Particle, sphg_particle, cd);
SPHG_SAME(Particle, sphg_particle, poro_cop); SPHG_SAME(Particle, sphg_particle, beta);
SPHG_SAME(Particle, sphg_particle, beta_ts); SPHG_SAME(Particle, sphg_particle, type);
SPHG_SAME(Particle, sphg_particle, matype); SPHG_SAME(Particle, sphg_particle, permtype);
SPHG_SAME(Particle, sphg_particle, etype);
static_assert(sizeof(Para_Soil) == sizeof(sphg_soil), "Para_Soil (Class_ParticleProperty.h:33-101) changed size");
SPHG_SAME(Para_Soil, sphg_soil, e); SPHG_SAME(Para_Soil, sphg_soil, c); SPHG_SAME(Para_Soil, sphg_soil, damp_coe);
static_assert(sizeof(Para_Fluid) == sizeof(sphg_fluid), "Para_Fluid (Class_ParticleProperty.h:104-133) changed size");
SPHG_SAME(Para_Fluid, sphg_fluid, bhtype); SPHG_SAME(Para_Fluid, sphg_fluid, strain_min);
static_assert(sizeof(Para_Structure) == sizeof(sphg_structure), "Para_Structure changed size");
SPHG_SAME(Para_Structure, sphg_structure, cop);
#undef SPHG_SAME

/* sphg_problem from the parameters SPH_Calculation holds after initial_pro / initial_mat / input_par
 * (CPP_1-SPHLoops.cpp:292-400): the flags, the problem controls, the free-surface thresholds (locals
 * of SPH_Calculation, CPP_1-SPHLoops.cpp:81-84), gravity, boundary, interaction and soil-structure
 * parameters, the model flags and the Cell_Con domain box and phase extents. */
inline sphg_problem make_problem(const Para_FL &fl, const Para_Pro &pp, double threshold_min, double threshold_max,
                                 const Para_GF &gf, const Para_Boundary &bo, const Para_Inter &in,
                                 const Para_SSInter &ss, const Para_Model &mo, const Cell_Con &cc) {
    sphg_problem p = {};
    p.fla = fl.fla; p.fls = fl.fls; p.flr = fl.flr; p.fli = fl.fli; p.flv = fl.flv; p.fleob = fl.fleob;
    p.flts = fl.flts; p.flspv = fl.flspv; p.flknl = fl.flknl; p.ntd = fl.ntd; p.flbndy = fl.flbndy;
    p.ndim = pp.ndim; p.fop = pp.fop; p.inip = pp.inip; p.loop = pp.loop; p.ntotal = pp.ntotal;
    p.step_regu = pp.step_regu; p.dttype = pp.dttype;
    p.dr = pp.dr; p.dt = pp.dt; p.tt = pp.tt;
    for (int k = 0; k < 3; k++) {
        p.xsph[k] = pp.xsph[k];
        p.art_vis[k][0] = pp.art_vis[k][0];
        p.art_vis[k][1] = pp.art_vis[k][1];
        p.move_vx[k] = bo.move_vx[k];
    }
    p.art_stre_coe = pp.art_stre_coe;
    p.threshold_min = threshold_min; p.threshold_max = threshold_max;
    p.gx = gf.gx; p.gy = gf.gy; p.gz = gf.gz;
    p.if_cal = bo.if_cal; p.if_move = bo.if_move; p.coe_bndy = bo.coe_bndy;
    p.interc = in.interc; p.permtype = in.permtype; p.pers = in.pers; p.perl = in.perl; p.perm = in.perm;
    p.ss_type = ss.type; p.ss_myu = ss.myu; p.ss_zeta = ss.zeta; p.ss_acc_type = ss.acc_type;
    p.flagem = mo.flagem; p.flagsb = mo.flagsb; p.flagdp = mo.flagdp; p.flagum = mo.flagum; p.flagam = mo.flagam;
    p.flagbui = mo.flagbui; p.flagpl = mo.flagpl; p.flagve = mo.flagve; p.flagfocc = mo.flagfocc;
    p.flagep = mo.flagep; p.alphats = mo.alphats;
    p.xmin = cc.xmin; p.xmax = cc.xmax; p.ymin = cc.ymin; p.ymax = cc.ymax; p.zmin = cc.zmin; p.zmax = cc.zmax;
    for (int k = 0; k < 3; k++) {
        p.soil_min[k] = cc.soil_min[k]; p.soil_max[k] = cc.soil_max[k];
        p.water_min[k] = cc.water_min[k]; p.water_max[k] = cc.water_max[k];
    }
    return p;
}

/* sphg_materials from the tables initial_mat fills (pSoil[30], pWater[3], pAir, pStructure;
 * CPP_2-initialf.cpp:325-852): member-for-member copies of layout-identical records. */
inline sphg_materials make_materials(const Para_Soil *soil, const Para_Fluid *water, const Para_Fluid &air,
                                     const Para_Structure &structure) {
    sphg_materials m;
    for (int k = 0; k < SPHG_NSOIL; k++) m.soil[k] = *reinterpret_cast<const sphg_soil *>(&soil[k]);
    for (int k = 0; k < 3; k++) m.water[k] = *reinterpret_cast<const sphg_fluid *>(&water[k]);
    m.air = *reinterpret_cast<const sphg_fluid *>(&air);
    m.structure = *reinterpret_cast<const sphg_structure *>(&structure);
    return m;
}

/* The particle array and the per-particle material copies a random field varies
 * (pParti_ConsPara[i].fai / c / cop / ds, CPP_17-1_K-L_expansion_non.cpp:575-600) to the device;
 * sphg_upload_done follows. */
inline int upload_particles(sphg_ctx *gpu, const Particle *par, const Para_Soil *consp, int n) {
    int rc = sphg_upload_particles(gpu, reinterpret_cast<const sphg_particle *>(par), n);
    if (rc) return rc;
    double *v = new double[n];
    const struct { const char *name; std::size_t off; } props[] = {
        {"ps_fai", offsetof(Para_Soil, fai)}, {"ps_c", offsetof(Para_Soil, c)},
        {"ps_cop", offsetof(Para_Soil, cop)}, {"ps_ds", offsetof(Para_Soil, ds)}};
    for (const auto &q : props) {
        for (int i = 0; i < n; i++) v[i] = *reinterpret_cast<const double *>(reinterpret_cast<const char *>(&consp[i]) + q.off);
        rc = sphg_set_field(gpu, q.name, v, n, 1);
        if (rc) break;
    }
    delete[] v;
    return rc;
}

/* The device state back into the particle array before the reference's writers run. */
inline int download_particles(sphg_ctx *gpu, Particle *par, int n, unsigned mask = SPHG_PART_ALL) {
    return sphg_download_particles(gpu, reinterpret_cast<sphg_particle *>(par), n, mask);
}

}  // namespace sphg_ref

#endif /* SPHG_REFERENCE_HPP_INCLUDED */
