// Particle-record upload and download (sphg_upload_particles / sphg_download_particles): the
// reference's array of Particle records (Class_particle.h:13-103) to and from the per-field
// device arrays, one member at a time through sphg_set_field / sphg_get_field.
#include <cstddef>
#include <cstring>
#include <vector>

#include "../../include/sphg.h"

namespace {

struct Member {
    const char *name;
    size_t off;
    int ncomp;
    unsigned group;
    bool upload;  // the device keeps it as state (weps, beta_ts and x are derived on download)
};

#define M(field, nc, grp, up) {#field, offsetof(sphg_particle, field), nc, grp, up}
const Member kMembers[] = {
    M(x, 3, SPHG_PART_KINEMATICS, false), M(xp, 3, SPHG_PART_KINEMATICS, true),
    M(vx, 3, SPHG_PART_KINEMATICS, true), M(vxp, 3, SPHG_PART_KINEMATICS, true),
    M(ax, 3, SPHG_PART_KINEMATICS, true), M(axp, 3, SPHG_PART_KINEMATICS, true),
    M(ux, 3, SPHG_PART_KINEMATICS, true),
    M(strep, 3, SPHG_PART_STRESS, true), M(interf, 3, SPHG_PART_STATE, true),
    M(interfss, 3, SPHG_PART_STATE, true),
    M(rho, 1, SPHG_PART_DENSITY, true), M(rhop, 1, SPHG_PART_DENSITY, true), M(mass, 1, SPHG_PART_DENSITY, true),
    M(hl, 1, SPHG_PART_DENSITY, true), M(pre, 1, SPHG_PART_DENSITY, true), M(prep, 1, SPHG_PART_DENSITY, true),
    M(e, 1, SPHG_PART_DENSITY, true),
    M(eps, 6, SPHG_PART_STRESS, true), M(deps, 6, SPHG_PART_STRESS, true), M(sig, 6, SPHG_PART_STRESS, true),
    M(weps, 9, SPHG_PART_STRESS, false), M(adps, 6, SPHG_PART_STRESS, true), M(tssig, 6, SPHG_PART_STRESS, true),
    M(tsqc, 1, SPHG_PART_STRESS, true),
    M(satu, 1, SPHG_PART_STATE, true), M(dsatu, 1, SPHG_PART_STATE, true), M(suct, 1, SPHG_PART_STATE, true),
    M(cop, 1, SPHG_PART_STATE, true),
    M(vsig, 1, SPHG_PART_STRESS, true), M(veps, 1, SPHG_PART_STRESS, true), M(divq, 1, SPHG_PART_STRESS, true),
    M(divr, 1, SPHG_PART_STRESS, true), M(meps, 1, SPHG_PART_STRESS, true),
    M(roue, 1, SPHG_PART_STATE, true), M(prew, 1, SPHG_PART_DENSITY, true), M(prea, 1, SPHG_PART_DENSITY, true),
    M(porosity, 1, SPHG_PART_DENSITY, true), M(cd, 1, SPHG_PART_STATE, true),
    M(poro_cop, 1, SPHG_PART_DENSITY, true),
    M(beta, 9, SPHG_PART_STRESS, true), M(beta_ts, 9, SPHG_PART_STRESS, false),
};
#undef M

const struct { const char *name; size_t off; } kIntMembers[] = {
    {"type", offsetof(sphg_particle, type)}, {"matype", offsetof(sphg_particle, matype)},
    {"permtype", offsetof(sphg_particle, permtype)}, {"etype", offsetof(sphg_particle, etype)}};

bool optional(const char *name) { return !strcmp(name, "tssig") || !strcmp(name, "tsqc"); }

}  // namespace

extern "C" int sphg_upload_particles(sphg_ctx *ctx, const sphg_particle *p, int n) {
    if (!ctx || !p || n <= 0) return SPHG_E_ARG;
    const char *base = reinterpret_cast<const char *>(p);
    std::vector<int> iv((size_t)n);
    for (const auto &m : kIntMembers) {
        for (int i = 0; i < n; i++) memcpy(&iv[i], base + (size_t)i * sizeof(sphg_particle) + m.off, sizeof(int));
        const int rc = sphg_set_field_i(ctx, m.name, iv.data(), n);
        if (rc) return rc;
    }
    std::vector<double> v;
    for (const auto &m : kMembers) {
        if (!m.upload) continue;
        v.resize((size_t)n * m.ncomp);
        for (int i = 0; i < n; i++)
            memcpy(&v[(size_t)i * m.ncomp], base + (size_t)i * sizeof(sphg_particle) + m.off, 8 * m.ncomp);
        const int rc = sphg_set_field(ctx, m.name, v.data(), n, m.ncomp);
        if (rc && !(rc == SPHG_E_ARG && optional(m.name))) return rc;
    }
    return SPHG_OK;
}

extern "C" int sphg_download_particles(sphg_ctx *ctx, sphg_particle *p, int n, unsigned mask) {
    if (!ctx || !p || n <= 0 || (mask & ~(unsigned)SPHG_PART_ALL)) return SPHG_E_ARG;
    char *base = reinterpret_cast<char *>(p);
    if (mask & SPHG_PART_IDS) {
        std::vector<int> iv((size_t)n);
        for (const auto &m : kIntMembers) {
            const int rc = sphg_get_field_i(ctx, m.name, iv.data(), n);
            if (rc) return rc;
            for (int i = 0; i < n; i++) memcpy(base + (size_t)i * sizeof(sphg_particle) + m.off, &iv[i], sizeof(int));
        }
    }
    std::vector<double> v;
    for (const auto &m : kMembers) {
        if (!(m.group & mask)) continue;
        v.resize((size_t)n * m.ncomp);
        const int rc = sphg_get_field(ctx, m.name, v.data(), n, m.ncomp);
        if (rc == SPHG_E_ARG && optional(m.name)) continue;
        if (rc) return rc;
        for (int i = 0; i < n; i++)
            memcpy(base + (size_t)i * sizeof(sphg_particle) + m.off, &v[(size_t)i * m.ncomp], 8 * m.ncomp);
    }
    return SPHG_OK;
}
