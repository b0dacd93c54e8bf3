// Host-side helpers of REBOUND's public interface that sit beside the
// force-and-step path: orbital elements, particle and simulation arithmetic,
// bulk particle access and the name hash.  The reference's Python package
// binds these through ctypes (rebound/particle.py, rebound/simulation.py,
// rebound/tools.py, rebound/units.py), so they keep the reference's
// signatures and arithmetic.
#include <cmath>
#include <cstring>

#include "internal.h"

namespace {

// tools.c:938,1027-1032: thresholds of the orbit calculation
constexpr double kTiny = 1.e-308;
constexpr double kMinInc = 1.e-8;
constexpr double kMinEcc = 1.e-8;

// acos(num/denom) in (-pi, pi]: negative when the disambiguator is; 0 or pi
// when |num/denom| reaches 1 (tools.c:1040-1053)
double acos_signed(double num, double denom, double disambiguator) {
    const double c = num / denom;
    if (c > -1. && c < 1.) {
        const double v = acos(c);
        return disambiguator < 0. ? -v : v;
    }
    return c <= -1. ? M_PI : 0.;
}

struct reb_orbit orbit_nan() {
    struct reb_orbit o;
    memset(&o, 0, sizeof(o));
    const double q = nan("");
    o.d = o.v = o.h = o.P = o.n = o.a = o.e = o.inc = q;
    o.Omega = o.omega = o.pomega = o.f = o.M = o.l = o.theta = o.T = o.rhill = q;
    return o;
}

}  // namespace

extern "C" {

// tools.c:48-56 (djb2): the hash the Python package stores unit names as
REB_API uint32_t reb_hash(const char* str) {
    uint32_t h = 5381;
    for (; *str; str++) h = h * 33u + (uint32_t)(unsigned char)*str;
    return h;
}

// The reference sets the OpenMP thread count of its CPU loops.  The force and
// step run on the GPU here; the host's collision resolve sizes its thread pool
// from the CPU count.  Accepted and ignored.
REB_API void reb_omp_set_num_threads(int num_threads) { (void)num_threads; }

// ------------------------------------------------------------ angles
// tools.c:504-507
REB_API double reb_mod2pi(double f) {
    const double pi2 = 2. * M_PI;
    return fmod(pi2 + fmod(f, pi2), pi2);
}

// tools.c:509-546: Kepler's equation by Newton's method, elliptic (initial
// guess of Danby & Burkardt) and hyperbolic
REB_API double reb_M_to_E(double e, double M) {
    double E;
    if (e < 1.) {
        M = reb_mod2pi(M);
        E = M + (M > M_PI ? -1. : 1.) * 0.71 * e;
        double F = E - e * sin(E) - M;
        for (int i = 0; i < 100; i++) {
            E = E - F / (1. - e * cos(E));
            F = E - e * sin(E) - M;
            if (fabs(F) < 1.e-15) break;
        }
        return reb_mod2pi(E);
    }
    E = M / fabs(M) * log(2. * fabs(M) / e + 1.8);
    double F = E - e * sinh(E) + M;
    for (int i = 0; i < 100; i++) {
        E = E - F / (1.0 - e * cosh(E));
        F = E - e * sinh(E) + M;
        if (fabs(F) < 1.e-15) break;
    }
    return E;
}

// tools.c:548-555
REB_API double reb_E_to_f(double e, double E) {
    if (e > 1.) return reb_mod2pi(2. * atan(sqrt((1. + e) / (e - 1.)) * tanh(0.5 * E)));
    return reb_mod2pi(2. * atan(sqrt((1. + e) / (1. - e)) * tan(0.5 * E)));
}

// tools.c:557-560
REB_API double reb_M_to_f(double e, double M) { return reb_E_to_f(e, reb_M_to_E(e, M)); }

// tools.c:1571-1583
REB_API struct reb_vec3d reb_tools_spherical_to_xyz(const double magnitude, const double theta, const double phi) {
    struct reb_vec3d v;
    v.x = magnitude * sin(theta) * cos(phi);
    v.y = magnitude * sin(theta) * sin(phi);
    v.z = magnitude * cos(theta);
    return v;
}

REB_API void reb_tools_xyz_to_spherical(const struct reb_vec3d xyz, double* magnitude, double* theta, double* phi) {
    *magnitude = sqrt(xyz.x * xyz.x + xyz.y * xyz.y + xyz.z * xyz.z);
    *theta = acos_signed(xyz.z, *magnitude, 1.);
    *phi = atan2(xyz.y, xyz.x);
}

// ------------------------------------------------------------ orbital elements
// tools.c:1055-1208: the elements of p about `primary`.  err: 1 the primary
// has no mass, 2 p sits on the primary (the elements are then NaN).
REB_API struct reb_orbit reb_orbit_from_particle_err(double G, struct reb_particle p, struct reb_particle primary,
                                                     int* err) {
    if (primary.m <= kTiny) {
        *err = 1;
        return orbit_nan();
    }
    struct reb_orbit o;
    memset(&o, 0, sizeof(o));
    const double mu = G * (p.m + primary.m);
    const double dx = p.x - primary.x, dy = p.y - primary.y, dz = p.z - primary.z;
    const double dvx = p.vx - primary.vx, dvy = p.vy - primary.vy, dvz = p.vz - primary.vz;
    o.d = sqrt(dx * dx + dy * dy + dz * dz);
    const double v2 = dvx * dvx + dvy * dvy + dvz * dvz;
    o.v = sqrt(v2);
    const double vcirc2 = mu / o.d;
    o.a = -mu / (v2 - 2. * vcirc2);
    o.rhill = o.a * cbrt(p.m / (3. * primary.m));
    // specific angular momentum
    const double hx = dy * dvz - dz * dvy;
    const double hy = dz * dvx - dx * dvz;
    const double hz = dx * dvy - dy * dvx;
    o.h = sqrt(hx * hx + hy * hy + hz * hz);
    o.hvec.x = hx;
    o.hvec.y = hy;
    o.hvec.z = hz;
    const double vdiff2 = v2 - vcirc2;
    if (o.d <= kTiny) {
        *err = 2;
        return orbit_nan();
    }
    const double vr = (dx * dvx + dy * dvy + dz * dvz) / o.d;
    const double rvr = o.d * vr;
    const double muinv = 1. / mu;
    // eccentricity vector
    const double ex = muinv * (vdiff2 * dx - rvr * dvx);
    const double ey = muinv * (vdiff2 * dy - rvr * dvy);
    const double ez = muinv * (vdiff2 * dz - rvr * dvz);
    o.e = sqrt(ex * ex + ey * ey + ez * ez);
    o.evec.x = ex;
    o.evec.y = ey;
    o.evec.z = ez;
    o.n = o.a / fabs(o.a) * sqrt(fabs(mu / (o.a * o.a * o.a)));   // negative when hyperbolic
    o.P = 2 * M_PI / o.n;
    o.inc = acos_signed(hz, o.h, 1.);
    // ascending node: z x h
    const double nx = -hy, ny = hx;
    const double n = sqrt(nx * nx + ny * ny);
    o.Omega = acos_signed(nx, n, ny);
    double ea;
    if (o.e < 1.) {
        ea = acos_signed(1. - o.d / o.a, o.e, vr);
        o.M = ea - o.e * sin(ea);
    } else {
        ea = acosh((1. - o.d / o.a) / o.e);
        if (vr < 0.) ea = -ea;
        o.M = o.e * sinh(ea) - ea;
    }
    const bool prograde = o.inc < M_PI / 2.;
    if (o.inc < kMinInc || o.inc > M_PI - kMinInc) {
        // nearly planar: longitudes from the vectors, angles from them
        o.theta = acos_signed(dx, o.d, dy);
        o.pomega = acos_signed(ex, o.e, ey);
        if (prograde) {
            o.omega = o.pomega - o.Omega;
            o.f = o.theta - o.pomega;
            o.l = o.e > kMinEcc ? o.pomega + o.M : o.theta - 2. * o.e * sin(o.f);
        } else {
            o.omega = o.Omega - o.pomega;
            o.f = o.pomega - o.theta;
            o.l = o.e > kMinEcc ? o.pomega - o.M : o.theta + 2. * o.e * sin(o.f);
        }
    } else {
        const double wpf = acos_signed(nx * dx + ny * dy, n * o.d, dz);   // omega + f
        o.omega = acos_signed(nx * ex + ny * ey, n * o.e, ez);
        if (prograde) {
            o.pomega = o.Omega + o.omega;
            o.f = wpf - o.omega;
            o.theta = o.Omega + wpf;
            o.l = o.e > kMinEcc ? o.pomega + o.M : o.theta - 2. * o.e * sin(o.f);
        } else {
            o.pomega = o.Omega - o.omega;
            o.f = wpf - o.omega;
            o.theta = o.Omega - wpf;
            o.l = o.e > kMinEcc ? o.pomega - o.M : o.theta + 2. * o.e * sin(o.f);
        }
    }
    const double t0 = p.sim ? p.sim->t : 0.0;
    o.T = t0 - o.M / fabs(o.n);
    o.f = reb_mod2pi(o.f);
    o.l = reb_mod2pi(o.l);
    o.M = reb_mod2pi(o.M);
    o.theta = reb_mod2pi(o.theta);
    o.omega = reb_mod2pi(o.omega);
    // Pal (2009) components of eccentricity and inclination
    const double fac = sqrt(2. / (1. + hz / o.h)) / o.h;
    o.pal_ix = -fac * hy;
    o.pal_iy = fac * hx;
    o.pal_k = o.h / mu * (dvy - dvz / (o.h + hz) * hy) - 1. / o.d * (dx - dz / (o.h + hz) * hx);
    o.pal_h = o.h / mu * (-dvx + dvz / (o.h + hz) * hx) - 1. / o.d * (dy - dz / (o.h + hz) * hy);
    return o;
}

REB_API struct reb_orbit reb_orbit_from_particle(double G, struct reb_particle p, struct reb_particle primary) {
    int err = 0;
    return reb_orbit_from_particle_err(G, p, primary, &err);
}

// ------------------------------------------------------------ particles
// particle.c:128-136: index of p in its simulation, -1 when it is not one of
// the simulation's particles
REB_API int reb_simulation_particle_index(struct reb_particle* p) {
    if (!p || !p->sim || p < p->sim->particles) return -1;
    const size_t i = (size_t)(p - p->sim->particles);
    return i < p->sim->N ? (int)i : -1;
}

// particle.c:137-145: the same for a variational particle
REB_API int reb_simulation_particle_var_index(struct reb_particle* p) {
    if (!p || !p->sim || p < p->sim->particles_var) return -1;
    const size_t i = (size_t)(p - p->sim->particles_var);
    return i < p->sim->N_var ? (int)i : -1;
}

// tools.c:454-461: centre of mass of the particles before p
REB_API struct reb_particle reb_simulation_jacobi_com(struct reb_particle* p) {
    const int i = reb_simulation_particle_index(p);
    if (i < 0) return reb_particle_nan();
    return reb_simulation_com_range(p->sim, 0, (size_t)i);
}

// particle.c:375-410
REB_API void reb_particle_isub(struct reb_particle* p1, struct reb_particle* p2) {
    p1->x -= p2->x; p1->y -= p2->y; p1->z -= p2->z;
    p1->vx -= p2->vx; p1->vy -= p2->vy; p1->vz -= p2->vz;
    p1->m -= p2->m;
}

REB_API void reb_particle_iadd(struct reb_particle* p1, struct reb_particle* p2) {
    p1->x += p2->x; p1->y += p2->y; p1->z += p2->z;
    p1->vx += p2->vx; p1->vy += p2->vy; p1->vz += p2->vz;
    p1->m += p2->m;
}

REB_API void reb_particle_imul(struct reb_particle* p1, double value) {
    p1->x *= value; p1->y *= value; p1->z *= value;
    p1->vx *= value; p1->vy *= value; p1->vz *= value;
    p1->m *= value;
}

REB_API double reb_particle_distance(struct reb_particle* p1, struct reb_particle* p2) {
    const double dx = p1->x - p2->x, dy = p1->y - p2->y, dz = p1->z - p2->z;
    return sqrt(dx * dx + dy * dy + dz * dz);
}

// ------------------------------------------------------------ simulations
// The host copy of the bodies must be current before it is read or changed
// here; changes are announced to the device afterwards.

// tools.c:1524-1535
REB_API void reb_simulation_imul(struct reb_simulation* r, double scalar_pos, double scalar_vel) {
    if (!rbamd::ensure_host(r)) return;
    for (size_t i = 0; i < r->N; i++) {
        struct reb_particle& p = r->particles[i];
        p.x *= scalar_pos; p.y *= scalar_pos; p.z *= scalar_pos;
        p.vx *= scalar_vel; p.vy *= scalar_vel; p.vz *= scalar_vel;
    }
    rbamd::mark_host_newer(r);
}

// tools.c:1537-1569: -1 when the particle counts differ
static int add_scaled(struct reb_simulation* r, struct reb_simulation* r2, double sign) {
    if (r->N != r2->N) return -1;
    if (!rbamd::ensure_host(r) || !rbamd::ensure_host(r2)) return -1;
    for (size_t i = 0; i < r->N; i++) {
        struct reb_particle& p = r->particles[i];
        const struct reb_particle& q = r2->particles[i];
        if (sign > 0.) {
            p.x += q.x; p.y += q.y; p.z += q.z;
            p.vx += q.vx; p.vy += q.vy; p.vz += q.vz;
        } else {
            p.x -= q.x; p.y -= q.y; p.z -= q.z;
            p.vx -= q.vx; p.vy -= q.vy; p.vz -= q.vz;
        }
    }
    rbamd::mark_host_newer(r);
    return 0;
}

REB_API int reb_simulation_iadd(struct reb_simulation* r, struct reb_simulation* r2) { return add_scaled(r, r2, 1.); }
REB_API int reb_simulation_isub(struct reb_simulation* r, struct reb_simulation* r2) { return add_scaled(r, r2, -1.); }

// simulation.c:801-861: bulk copies of masses, radii, positions and
// velocities (any pointer may be NULL)
REB_API void reb_simulation_get_serialized_particle_data(struct reb_simulation* r, double* m, double* radius,
                                                         double (*xyz)[3], double (*vxvyvz)[3],
                                                         double (*xyzvxvyvz)[6]) {
    if (!rbamd::ensure_host(r)) return;
    for (size_t i = 0; i < r->N; i++) {
        const struct reb_particle& p = r->particles[i];
        if (m) m[i] = p.m;
        if (radius) radius[i] = p.r;
        if (xyz) {
            xyz[i][0] = p.x; xyz[i][1] = p.y; xyz[i][2] = p.z;
        }
        if (vxvyvz) {
            vxvyvz[i][0] = p.vx; vxvyvz[i][1] = p.vy; vxvyvz[i][2] = p.vz;
        }
        if (xyzvxvyvz) {
            xyzvxvyvz[i][0] = p.x; xyzvxvyvz[i][1] = p.y; xyzvxvyvz[i][2] = p.z;
            xyzvxvyvz[i][3] = p.vx; xyzvxvyvz[i][4] = p.vy; xyzvxvyvz[i][5] = p.vz;
        }
    }
}

REB_API void reb_simulation_set_serialized_particle_data(struct reb_simulation* r, double* m, double* radius,
                                                         double (*xyz)[3], double (*vxvyvz)[3],
                                                         double (*xyzvxvyvz)[6]) {
    if (!rbamd::ensure_host(r)) return;
    for (size_t i = 0; i < r->N; i++) {
        struct reb_particle& p = r->particles[i];
        if (m) p.m = m[i];
        if (radius) p.r = radius[i];
        if (xyz) {
            p.x = xyz[i][0]; p.y = xyz[i][1]; p.z = xyz[i][2];
        }
        if (vxvyvz) {
            p.vx = vxvyvz[i][0]; p.vy = vxvyvz[i][1]; p.vz = vxvyvz[i][2];
        }
        if (xyzvxvyvz) {
            p.x = xyzvxvyvz[i][0]; p.y = xyzvxvyvz[i][1]; p.z = xyzvxvyvz[i][2];
            p.vx = xyzvxvyvz[i][3]; p.vy = xyzvxvyvz[i][4]; p.vz = xyzvxvyvz[i][5];
        }
    }
    r->did_modify_particles = 1;
    rbamd::mark_host_newer(r);
}

}  // extern "C"
