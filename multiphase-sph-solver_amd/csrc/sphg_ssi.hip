// Soil-structure interaction on the device: interaction_ss_rigid / _deformable / _penalty
// (CPP_10-1-interact_ss.cpp:17-1038) with the file's geometry helpers (:1039-1369).
//
// The reference runs two loops: per soil / water particle it finds the contact on the line (2D)
// or triangle (3D) of its nearest structure particles, sets its own contact force and scatters
// the reaction onto those structure particles (+= into per-thread slots); a second loop turns
// each particle's slot into interfss and adds interfss / mass to ax. Here:
//   k_ssi_soil    one thread per soil / water particle: contact geometry and force, own interfss
//                 and ax, and the reaction parts for up to three structure particles, kept in
//                 ssi_id / ssi_c instead of scattered;
//   k_ssi_struct  one thread per structure particle: gathers the reaction parts addressed to it
//                 from its soil / water neighbours in ascending particle order, the order of the
//                 reference's serial += sequence (one thread), then interfss and ax.
//   k_accel_tail  the damping clamp, the free-solid wall contact (flbndy 5, boundary_free_solid,
//                 CPP_13-boundary.cpp:1239-1515) and the velocity update that end the momentum pass.
// All expressions keep the reference's operation order (FP64, contraction off). Quirks kept:
// dist_sorting swaps ids but not distances; clLine's C and clPlane's D keep their constructor
// value 1; the mapping point is accepted by |distance| < 1e-3 in absolute units
// (cal_mappoint_line / _plane). One deviation: where set_normal() finds |(A, B(, C))| <= 1e-10
// the reference keeps the OpenMP-private normal of the thread's previous contact; a thread here
// has no previous contact and keeps (0, 0, 0) (such a line needs structure particles about 1e10
// length units apart).
#include <hip/hip_runtime.h>

namespace sphg {

struct SPt {
    double x, y, z;
};

__device__ __forceinline__ SPt spt_of(const State &S, int cur, int j) {
    return SPt{S.xp[cur][0][j], S.xp[cur][1][j], S.xp[cur][2][j]};
}

// Vector3 dot product (Vector3.h:101-104)
__device__ __forceinline__ double sdot(SPt a, SPt b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ SPt ssub(SPt a, SPt b) { return SPt{a.x - b.x, a.y - b.y, a.z - b.z}; }

__device__ __forceinline__ double pt_dist(SPt a, SPt b) {
    return sqrt((b.x - a.x) * (b.x - a.x) + (b.y - a.y) * (b.y - a.y) + (b.z - a.z) * (b.z - a.z));
}

// cal_dpline / cal_dpplane (:1224-1250); the line's C and the plane's D are 1
__device__ __forceinline__ double dp_line(SPt p, double A, double B) {
    const double dst = fabs(A * p.x + B * p.y + 1.0), temp = sqrt(A * A + B * B);
    return temp > 1.0e-3 ? dst / temp : 0.0;
}
__device__ __forceinline__ double dp_plane(SPt p, double A, double B, double C) {
    const double dst = fabs(A * p.x + B * p.y + C * p.z + 1.0), temp = sqrt(A * A + B * B + C * C);
    return temp > 1.0e-3 ? dst / temp : 0.0;
}

// linear_interpolate (:1187-1202)
__device__ __forceinline__ void ss_interp(SPt p1, SPt p2, SPt r, double m1, double m2, const double v1[3],
                                          const double v2[3], double &m, double v[3]) {
    const double xk1 = 0.0, xk2 = pt_dist(p1, p2), xp = pt_dist(p1, r);
    m = (m1 * fabs(xp - xk2) + m2 * fabs(xp - xk1)) / fabs(xk2 - xk1);
#pragma unroll
    for (int k = 0; k < 3; k++) v[k] = (v1[k] * fabs(xp - xk2) + v2[k] * fabs(xp - xk1)) / fabs(xk2 - xk1);
}

// linear_interpolate_force (:1204-1222)
__device__ __forceinline__ void ss_interp_force(SPt p1, SPt p2, SPt r, const double f[3], double o1[3],
                                                double o2[3]) {
    const double xk1 = 0.0, xk2 = pt_dist(p1, p2), xp = pt_dist(p1, r);
#pragma unroll
    for (int k = 0; k < 3; k++) o1[k] = -fabs(xp - xk2) / fabs(xk1 - xk2) * f[k];
#pragma unroll
    for (int k = 0; k < 3; k++) o2[k] = -fabs(xp - xk1) / fabs(xk1 - xk2) * f[k];
}

// Contact of particle i with the line (2D) / triangle (3D) through its nearest particles of type
// `wall` (structure particles for the soil-structure interaction, boundary particles for the
// free-solid wall), as interaction_ss_* and boundary_free_solid compute it: on = contact found
// within `reach` of the line / plane, the mapped point inside the segment / triangle and, in 3D,
// the two lines intersecting; ms / vs the wall mass and velocity interpolated at the contact point
// from the walls' vxp (or their interfss, use_interfss).
struct Contact {
    bool on;
    int id[3];
    SPt p[3], outp, sect;
    double dp, n[3], ms, vs[3];
};

__device__ __forceinline__ void wall_vel(const State &S, int j, bool use_interfss, double v[3]) {
#pragma unroll
    for (int k = 0; k < 3; k++) v[k] = use_interfss ? S.interfss[k][j] : S.vxp[k][j];
}

template <int NDIM>
__device__ Contact find_contact(const State &S, const Params &P, int cur, int i, int wall, bool use_interfss,
                                double reach) {
    Contact c;
    c.on = false;
    if (!(S.nb_types[i] & (1 << wall))) return c;  // no wall particle listed
    const SPt p4 = spt_of(S, cur, i);
    // wall neighbours in list order, their distances over all three coordinates, and the first
    // NDIM ids after clParti_Pair::dist_sorting (Class_particle.cpp:343-359), formed while
    // streaming: an arriving id passes the occupied slots in order and is exchanged with each
    // slot whose (unmoved) distance exceeds its own
    int ids[3] = {-1, -1, -1}, total = 0;
    double d0[3] = {0.0, 0.0, 0.0};
    const int nc = nb_count(S, i);
    for (int k = 0; k < nc; k++) {
        const int e = S.nbr[(size_t)k * P.cap + i];
        if (nb_type(e) != wall) continue;
        const int j = nb_id(e);
        const SPt pj = spt_of(S, cur, j);
        double d = (p4.x - pj.x) * (p4.x - pj.x) + (p4.y - pj.y) * (p4.y - pj.y) + (p4.z - pj.z) * (p4.z - pj.z);
        d = sqrt(d);
        int x = j;
        const int passes = total < NDIM ? total : NDIM;
        for (int q = 0; q < passes; q++)
            if (d0[q] > d) {
                const int t = ids[q];
                ids[q] = x;
                x = t;
            }
        if (total < NDIM) {
            ids[total] = x;
            d0[total] = d;
        }
        total++;
    }
    if (total < NDIM) return c;
#pragma unroll
    for (int q = 0; q < 3; q++) {
        c.id[q] = ids[q];
        c.p[q] = q < NDIM ? spt_of(S, cur, ids[q]) : SPt{0.0, 0.0, 0.0};
    }
    const SPt p1 = c.p[0], p2 = c.p[1], p3 = c.p[2];
    if (NDIM == 2) {
        const double den = p2.y * p1.x - p1.y * p2.x;  // cal_line (:1055-1061)
        const double A = (p1.y - p2.y) / den, B = (p2.x - p1.x) / den;
        const double dp = dp_line(p4, A, B);
        SPt outp;  // cal_mappoint_line (:1094-1125); the mapped point's z is 0
        {
            const double temp = sqrt(A * A + B * B);
            const SPt q1{p4.x + dp * A / temp, p4.y + dp * B / temp, 0.0};
            const SPt q2{p4.x - dp * A / temp, p4.y - dp * B / temp, 0.0};
            outp = dp_line(q1, A, B) < 1.0e-3 ? q1 : q2;
        }
        // check_ifinside_line (:1296-1315)
        const SPt v0 = ssub(p2, p1), v1 = ssub(SPt{outp.x, outp.y, 0.0}, p1);
        const double u = 1.0 / sdot(v0, v0) * sdot(v0, v1);
        if ((u < 0 || u > 1) || !(dp < reach)) return c;
        double nx = 0.0, ny = 0.0;
        const double t = sqrt(A * A + B * B);  // clLine::set_normal
        if (t > 1.0e-10) {
            nx = A / t;
            ny = B / t;
        }
        const SPt ev{outp.x - p4.x, outp.y - p4.y, 0.0}, nv{nx, ny, 0.0};
        if (!(sdot(ev, nv) <= 0.0)) {  // check_ifoutward2D
            nx = -nx;
            ny = -ny;
        }
        double v1v[3], v2v[3];
        wall_vel(S, ids[0], use_interfss, v1v);
        wall_vel(S, ids[1], use_interfss, v2v);
        ss_interp(p1, p2, outp, S.mass[ids[0]], S.mass[ids[1]], v1v, v2v, c.ms, c.vs);
        c.on = true;
        c.dp = dp;
        c.outp = outp;
        c.n[0] = nx;
        c.n[1] = ny;
        c.n[2] = 0.0;
        return c;
    }
    // cal_plane (:1063-1076)
    const double den = -p1.x * (p2.y * p3.z - p3.y * p2.z) - p2.x * (p3.y * p1.z - p1.y * p3.z) -
                       p3.x * (p1.y * p2.z - p2.y * p1.z);
    double A = p1.y * (p2.z - p3.z) + p2.y * (p3.z - p1.z) + p3.y * (p1.z - p2.z);
    double B = p1.z * (p2.x - p3.x) + p2.z * (p3.x - p1.x) + p3.z * (p1.x - p2.x);
    double C = p1.x * (p2.y - p3.y) + p2.x * (p3.y - p1.y) + p3.x * (p1.y - p2.y);
    A = A / den;
    B = B / den;
    C = C / den;
    const double dp = dp_plane(p4, A, B, C);
    SPt outp;  // cal_mappoint_plane (:1127-1160)
    {
        const double temp = sqrt(A * A + B * B + C * C);
        const SPt q1{p4.x + dp * A / temp, p4.y + dp * B / temp, p4.z + dp * C / temp};
        const SPt q2{p4.x - dp * A / temp, p4.y - dp * B / temp, p4.z - dp * C / temp};
        outp = dp_plane(q1, A, B, C) < 1.0e-3 ? q1 : q2;
    }
    // check_ifinside_triangle (:1317-1344)
    {
        const SPt v0 = ssub(p3, p1), v1 = ssub(p2, p1), v2 = ssub(outp, p1);
        const double d00 = sdot(v0, v0), d01 = sdot(v0, v1), d02 = sdot(v0, v2), d11 = sdot(v1, v1),
                     d12 = sdot(v1, v2);
        const double inv = 1 / (d00 * d11 - d01 * d01);
        const double u = (d11 * d02 - d01 * d12) * inv;
        if (u < 0 || u > 1) return c;
        const double v = (d00 * d12 - d01 * d02) * inv;
        if ((v < 0 || v > 1) || !(u + v <= 1)) return c;
    }
    if (!(dp < reach)) return c;
    double nx = 0.0, ny = 0.0, nz = 0.0;
    {
        const double t = sqrt(A * A + B * B + C * C);  // clPlane::set_normal
        if (t > 1.0e-10) {
            nx = A / t;
            ny = B / t;
            nz = C / t;
        }
        const SPt ev = ssub(outp, p4), nv{nx, ny, nz};
        if (!(sdot(ev, nv) <= 0.0)) {  // check_ifoutward3D
            nx = -nx;
            ny = -ny;
            nz = -nz;
        }
    }
    // cal_line3D (:1078-1092) for p1-p2 and p3-outp, cal_intersect_lines3D (:1162-1185)
    auto line3 = [](SPt a, SPt b, double dir[3]) {
        double dst = 0.0;
        dst = (b.x - a.x) * (b.x - a.x) + (b.y - a.y) * (b.y - a.y) + (b.z - a.z) * (b.z - a.z);
        dst = sqrt(dst);
        dir[0] = (b.x - a.x) / dst;
        dir[1] = (b.y - a.y) / dst;
        dir[2] = (b.z - a.z) / dst;
    };
    double l1[3], l2[3];
    line3(p1, p2, l1);
    line3(p3, outp, l2);
    const double flag = l1[0] * l2[1] - l2[0] * l1[1];
    if (!(fabs(flag) > 1.0e-3)) return c;
    const double t = (l2[0] * (p1.y - p3.y) - l2[1] * (p1.x - p3.x)) / flag;
    const SPt sect{p1.x + l1[0] * t, p1.y + l1[1] * t, p1.z + l1[2] * t};
    double v1v[3], v2v[3], v3v[3], tvs[3], tms;
    wall_vel(S, ids[0], use_interfss, v1v);
    wall_vel(S, ids[1], use_interfss, v2v);
    wall_vel(S, ids[2], use_interfss, v3v);
    ss_interp(p1, p2, sect, S.mass[ids[0]], S.mass[ids[1]], v1v, v2v, tms, tvs);
    ss_interp(p3, sect, outp, S.mass[ids[2]], tms, v3v, tvs, c.ms, c.vs);
    c.on = true;
    c.dp = dp;
    c.outp = outp;
    c.sect = sect;
    c.n[0] = nx;
    c.n[1] = ny;
    c.n[2] = nz;
    return c;
}

// Contact force of particle i (interaction_ss_* body): fn / ft from the relative velocity vs
// (contact-point velocity minus the particle's vxp), split along the normal, Coulomb cap myu.
// deformable: vs scaled by ms / (ms + m_i) (scale); penalty: normal force from the overlap.
template <int NDIM>
__device__ __forceinline__ void contact_force(const State &S, const Params &P, int i, const Contact &c, bool scale,
                                              bool penalty, double fn[3], double ft[3]) {
    const double dt = *S.dtp, dr = P.dr, myu = P.ss_myu, zeta = P.ss_zeta, mi = S.mass[i];
    double vs[3] = {c.vs[0], c.vs[1], c.vs[2]};
#pragma unroll
    for (int k = 0; k < NDIM; k++) vs[k] -= S.vxp[k][i];
    if (scale) {
#pragma unroll
        for (int k = 0; k < NDIM; k++) vs[k] *= c.ms / (c.ms + mi);
    }
    // decomposition_velocity2D / 3D (:1346-1369)
    double v = vs[0] * c.n[0] + vs[1] * c.n[1];
    if (NDIM == 3) v = v + vs[2] * c.n[2];
    double vn[3] = {0.0, 0.0, 0.0}, vt[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int k = 0; k < NDIM; k++) {
        vn[k] = v * c.n[k];
        vt[k] = vs[k] - vn[k];
    }
    if (penalty) {
        const double mag = (1.0 - zeta) * 2.0 * mi * (dr - c.dp) / dt / dt;
#pragma unroll
        for (int k = 0; k < NDIM; k++) fn[k] = mag * c.n[k];
    } else {
#pragma unroll
        for (int k = 0; k < NDIM; k++) fn[k] = mi * vn[k] / dt;
    }
#pragma unroll
    for (int k = 0; k < NDIM; k++) ft[k] = mi * vt[k] / dt;
    double mag_fn = fn[0] * fn[0] + fn[1] * fn[1], mag_ft = ft[0] * ft[0] + ft[1] * ft[1];
    if (NDIM == 3) {
        mag_fn = mag_fn + fn[2] * fn[2];
        mag_ft = mag_ft + ft[2] * ft[2];
    }
    mag_fn = sqrt(mag_fn);
    mag_ft = sqrt(mag_ft);
    if (mag_ft > (mag_fn * myu)) {
#pragma unroll
        for (int k = 0; k < NDIM; k++) vt[k] = vt[k] * myu * mag_fn / mag_ft;
#pragma unroll
        for (int k = 0; k < NDIM; k++) ft[k] = mi * vt[k] / dt;
    }
}

template <int NDIM>
__global__ __launch_bounds__(256) void k_ssi_soil(State S, Params P, int cur) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P.n) return;
    const int ti = S.type[i];
    if (ti != 1 && ti != 2) return;
    const int var = P.ss_type;
    const double zeta = P.ss_zeta, mi = S.mass[i];
    double fn[3] = {0.0, 0.0, 0.0}, ft[3] = {0.0, 0.0, 0.0};
    int tid[3] = {-1, -1, -1};
    double cn[3][3], ct[3][3];
#pragma unroll
    for (int q = 0; q < 3; q++)
#pragma unroll
        for (int k = 0; k < 3; k++) cn[q][k] = ct[q][k] = 0.0;
    const Contact c = find_contact<NDIM>(S, P, cur, i, 4, false, var == 1 ? 0.5 * P.dr : P.dr);
    if (c.on) {
        contact_force<NDIM>(S, P, i, c, var != 1, var == 3, fn, ft);
        if (NDIM == 2) {
            ss_interp_force(c.p[0], c.p[1], c.outp, fn, cn[0], cn[1]);
            ss_interp_force(c.p[0], c.p[1], c.outp, ft, ct[0], ct[1]);
            tid[0] = c.id[0];
            tid[1] = c.id[1];
        } else {
            double f3n[3], f3t[3], a1[3], a2[3];
            ss_interp_force(c.p[2], c.sect, c.outp, fn, cn[0], f3n);
            ss_interp_force(c.p[2], c.sect, c.outp, ft, ct[0], f3t);
            // id1 / id2 receive -= of the second split: kept as + of the negated parts
            ss_interp_force(c.p[0], c.p[1], c.sect, f3n, a1, a2);
#pragma unroll
            for (int k = 0; k < 3; k++) {
                cn[1][k] = -a1[k];
                cn[2][k] = -a2[k];
            }
            ss_interp_force(c.p[0], c.p[1], c.sect, f3t, a1, a2);
#pragma unroll
            for (int k = 0; k < 3; k++) {
                ct[1][k] = -a1[k];
                ct[2][k] = -a2[k];
            }
            tid[0] = c.id[2];
            tid[1] = c.id[0];
            tid[2] = c.id[1];
        }
    }
#pragma unroll
    for (int q = 0; q < 3; q++) {
        S.ssi_id[q][i] = tid[q];
#pragma unroll
        for (int k = 0; k < 3; k++) {
            S.ssi_c[q][k][i] = cn[q][k];
            S.ssi_c[q][3 + k][i] = ct[q][k];
        }
    }
    // the particle's own slot into interfss and ax
#pragma unroll
    for (int k = 0; k < 3; k++) {
        double f = 0.0;
        if (var == 3) f += ft[k] + fn[k];
        else f += zeta * (ft[k] + fn[k]);
        S.interfss[k][i] = f;
        S.ax[k][i] += f / mi;
    }
}

// reaction sums of the structure particles, ascending in the contributing particle's index
__global__ __launch_bounds__(256) void k_ssi_struct(State S, Params P) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P.n || S.type[i] != 4) return;
    const int var = P.ss_type;
    const double zeta = P.ss_zeta;
    double fn[3] = {0.0, 0.0, 0.0}, ft[3] = {0.0, 0.0, 0.0};
    const int nc = nb_count(S, i);
    int last = -1;
    for (;;) {
        int best = 0x7fffffff, bq = -1;
        for (int k = 0; k < nc; k++) {
            const int e = S.nbr[(size_t)k * P.cap + i];
            const int tj = nb_type(e), j = nb_id(e);
            if ((tj != 1 && tj != 2) || j <= last || j >= best) continue;
#pragma unroll
            for (int q = 0; q < 3; q++)
                if (S.ssi_id[q][j] == i) {
                    best = j;
                    bq = q;
                }
        }
        if (bq < 0) break;
#pragma unroll
        for (int q = 0; q < 3; q++) {
            if (q != bq) continue;
#pragma unroll
            for (int k = 0; k < 3; k++) {
                fn[k] += S.ssi_c[q][k][best];
                ft[k] += S.ssi_c[q][3 + k][best];
            }
        }
        last = best;
    }
    const double mi = S.mass[i];
#pragma unroll
    for (int k = 0; k < 3; k++) {
        double f = 0.0;
        if (var == 3) f += ft[k] + fn[k];
        else f += zeta * (ft[k] + fn[k]);
        S.interfss[k][i] = f;
        S.ax[k][i] += f / mi;
    }
}

// the rest of the momentum pass after the interaction sums (see k_accel)
template <int NDIM>
__global__ __launch_bounds__(256) void k_accel_tail(State S, Params P, int cur) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P.n) return;
    const int ti = S.type[i];
    double ax[3] = {S.ax[0][i], S.ax[1][i], S.ax[2][i]};
    const double vi[3] = {S.vxp[0][i], S.vxp[1][i], S.vxp[2][i]};
    const double axp[3] = {S.axp[0][i], S.axp[1][i], S.axp[2][i]};
    const bool damp_on = (ti == 2 || ti == 4) && step_of(S) <= P.inip;
    accel_damp<NDIM>(S, P, i, ax, vi, axp, damp_on);
    if (P.flv > 0 && (ti == 2 || ti == 4)) {
        // soil_damping_vibration's clamp (CPP_9-1-acceleration_vibration.cpp:207-221)
        double tempdown = S.vdamp[3][i];
        for (int k = 0; k < NDIM; k++) {
            const double damp = S.vdamp[k][i];
            if (fabs(tempdown) < 1.0e-7) tempdown = 1.0e-7;
            const double vx1 = vi[k] + 0.5 * (axp[k] + ax[k]) * *S.dtp;
            const double vx2 = vi[k] + 0.5 * (axp[k] + ax[k] - damp / tempdown) * *S.dtp;
            if (vx1 * vx2 <= 0.0) ax[k] = -2.0 * vi[k] / *S.dtp - axp[k];
            else ax[k] -= damp / tempdown;
        }
    }
    if (P.flbndy == 5 && (NDIM == 3 || (ti >= 1 && ti <= 4))) {
        // boundary_free_solid (CPP_13-boundary.cpp:1239-1515): the rigid contact against the wall
        // (boundary particles, their interfss as velocity); the 3D loop has no type test
        const Contact c = find_contact<NDIM>(S, P, cur, i, 0, true, 0.5 * P.dr);
        if (c.on) {
            double fn[3] = {0.0, 0.0, 0.0}, ft[3] = {0.0, 0.0, 0.0};
            contact_force<NDIM>(S, P, i, c, false, false, fn, ft);
            const double mi = S.mass[i], zeta = P.ss_zeta;
#pragma unroll
            for (int k = 0; k < NDIM; k++) ax[k] += zeta * (ft[k] + fn[k]) / mi;
            if (NDIM == 2) ax[2] += 0.0;
        }
    }
    if (P.flv > 0 && ti == 0) {
        // with vibration loads the walls move in k_vib_walls_vel, after this kernel
#pragma unroll
        for (int k = 0; k < 3; k++) S.ax[k][i] = ax[k];
        return;
    }
    accel_vel<NDIM>(S, P, i, ti, cur, ax, vi, axp);
}

template __global__ void k_ssi_soil<2>(State, Params, int);
template __global__ void k_ssi_soil<3>(State, Params, int);
template __global__ void k_accel_tail<2>(State, Params, int);
template __global__ void k_accel_tail<3>(State, Params, int);

}  // namespace sphg
