// Host driver of one project directory, the way the reference's main and SPH_Calculation run one
// (CPP_0-main.cpp, CPP_1-SPHLoops.cpp:67-782), for the flag set this library runs: read
// <dir>/input.dat, set up the initial particle state, run the time loop on the device and write
// <dir>/data-<step>.odb at the reference's output steps.
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/sphg.h"

namespace {

struct OutCtx {
    sphg_ctx *ctx;
    const char *dir;
    int n;
    int last;
};

int write_step(void *user, int step, double t) {
    OutCtx *o = static_cast<OutCtx *>(user);
    o->last = step;
    return sphg_write_odb(o->ctx, o->dir, step, t, 0, o->n);
}

int upload(sphg_ctx *ctx, int n, const std::vector<std::pair<const char *, std::pair<const double *, int>>> &d,
           const std::vector<std::pair<const char *, const int *>> &iv) {
    for (const auto &f : iv)
        if (int rc = sphg_set_field_i(ctx, f.first, f.second, n)) return rc;
    for (const auto &f : d)
        if (int rc = sphg_set_field(ctx, f.first, f.second.first, n, f.second.second)) return rc;
    return sphg_upload_done(ctx, n);
}

}  // namespace

extern "C" int sphg_solve_project(const char *dir, int precision, int *last_step) {
    if (!dir) return SPHG_E_ARG;
    if (last_step) *last_step = 0;
    const std::string inp = std::string(dir) + "/input.dat";
    sphg_problem pro;
    sphg_materials mat;
    int n = 0;
    if (int rc = sphg_read_input(inp.c_str(), &pro, &mat, &n, 0, nullptr, nullptr, nullptr, nullptr, nullptr,
                                 nullptr))
        return rc;
    // random fields come from the SF/SV sections, which the reader does not take; the K-L entry
    // points serve them through the API
    if (pro.flspv != 0) return SPHG_E_UNSUPPORTED;
    const size_t N = (size_t)n;
    std::vector<double> xp(3 * N), vxp(3 * N);
    std::vector<int> type(N), matype(N), permtype(N), etype(N);
    if (int rc = sphg_read_input(inp.c_str(), &pro, &mat, &n, n, xp.data(), vxp.data(), type.data(), matype.data(),
                                 permtype.data(), etype.data()))
        return rc;
    // setting_vis_cons + setting_par (CPP_1-SPHLoops.cpp:400, :440)
    std::vector<double> rho(N), rhop(N), mass(N), hl(N), satu(N), e(N), por(N), axp(3 * N), cd(N), roue(N),
        sig(6 * N), pre(N), prep(N), suct(N);
    if (int rc = sphg_setting_par(&pro, &mat, n, xp.data(), type.data(), matype.data(), etype.data(), rho.data(),
                                  rhop.data(), mass.data(), hl.data(), satu.data(), e.data(), por.data(), axp.data(),
                                  cd.data(), roue.data(), sig.data(), pre.data(), prep.data(), suct.data()))
        return rc;
    sphg_ctx *ctx = nullptr;
    if (int rc = sphg_create(&ctx, &pro, &mat, n, precision)) return rc;
    if (pro.flv >= 1) {
        // the VL section and vibra.dat (CPP_1-SPHLoops.cpp:306-340)
        double tt_, st_, et_;
        int ns = 0;
        int rv = sphg_read_vibration(dir, &tt_, &ns, &st_, &et_, nullptr, 0);
        std::vector<double> loads(3 * (size_t)(ns > 0 ? ns : 1));
        if (rv == SPHG_OK) rv = sphg_read_vibration(dir, &tt_, &ns, &st_, &et_, loads.data(), ns);
        if (rv == SPHG_OK) rv = sphg_set_vibration(ctx, tt_, ns, st_, et_, loads.data());
        if (rv) {
            sphg_destroy(ctx);
            return rv;
        }
    }
    if (pro.flr == 1) {
        // the RP section (CPP_2-initialf.cpp:180-201)
        sphg_rain r{};
        int rv = sphg_read_rain(dir, &r);
        if (rv == SPHG_OK) rv = sphg_set_rain(ctx, &r);
        if (rv) {
            sphg_destroy(ctx);
            return rv;
        }
    }
    if (pro.fleob == 1) {
        // the EC section (CPP_2-initialf.cpp:125-157); no stage step is given there (see the reader)
        int ec = 0;
        int rv = sphg_read_excavation(dir, &ec, nullptr, 0);
        std::vector<int> ecp(3 * (size_t)(ec > 0 ? ec : 1));
        if (rv == SPHG_OK) rv = sphg_read_excavation(dir, &ec, ecp.data(), ec);
        if (rv == SPHG_OK) rv = sphg_set_excavation(ctx, ec, ecp.data());
        if (rv) {
            sphg_destroy(ctx);
            return rv;
        }
    }
    int rc = upload(ctx, n,
                    {{"xp", {xp.data(), 3}}, {"vxp", {vxp.data(), 3}}, {"rho", {rho.data(), 1}},
                     {"rhop", {rhop.data(), 1}}, {"mass", {mass.data(), 1}}, {"hl", {hl.data(), 1}},
                     {"satu", {satu.data(), 1}}, {"e", {e.data(), 1}}, {"porosity", {por.data(), 1}},
                     {"axp", {axp.data(), 3}}, {"cd", {cd.data(), 1}}, {"roue", {roue.data(), 1}},
                     {"sig", {sig.data(), 6}}, {"pre", {pre.data(), 1}}, {"prep", {prep.data(), 1}},
                     {"suct", {suct.data(), 1}}},
                    {{"type", type.data()}, {"matype", matype.data()}, {"permtype", permtype.data()},
                     {"etype", etype.data()}});
    if (rc == SPHG_OK) rc = sphg_set_step(ctx, 1);
    OutCtx oc{ctx, dir, n, 0};
    if (rc == SPHG_OK) {
        double t = 0.0;
        rc = sphg_run(ctx, pro.loop, pro.fop, pro.tt, &t, write_step, &oc);
        // after a step error the reference writes the state it stopped in (CPP_1-SPHLoops.cpp:525-541);
        // here that is the state after the last step of the failing batch
        if (rc == SPHG_E_NOPARTICLES || rc == SPHG_E_CELLRANGE || rc == SPHG_E_NEIGHBOURS) {
            const int l = sphg_current_step(ctx) - 1;
            if (sphg_write_odb(ctx, dir, l, t, 0, n) == SPHG_OK) oc.last = l;
        }
    }
    if (last_step) *last_step = oc.last;
    sphg_destroy(ctx);
    return rc;
}
