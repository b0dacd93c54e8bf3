// Host side of the drop-in ABI: lifecycle, messages, integrator registry,
// particle bookkeeping and the reference's input generators.
//
// The generators (rand_r stream, Plummer sphere, orbital elements) must
// reproduce the reference's numbers bit for bit, so they evaluate the same
// expressions in the same order with the host's libm (glibc), as the
// reference does.  They run once at set-up; the force-and-step path itself
// runs on the GPU (gravity.hip, integrators.hip, whfast.hip, tree.hip).
#include <cmath>
#include <cstdarg>
#include <cstring>
#include <mutex>
#include <new>
#include <sys/time.h>
#include <unistd.h>

#include "internal.h"

using namespace rbamd;

extern "C" {

// ===================================================================== messages
// rebound.c:126-190.  Ten-slot ring when save_messages is set, else stderr.
static const size_t kMessagesMax = 10;

static void push_message(struct reb_simulation* r, char type, const char* msg) {
    if (!r || !r->save_messages) {
        const char* label = type == 'e' ? "Error!" : (type == 'w' ? "Warning!" : "REBOUND Message!");
        fprintf(stderr, "\n%s %s\n", label, msg);
        return;
    }
    if (!r->messages) r->messages = (char**)calloc(kMessagesMax, sizeof(char*));
    size_t n = 0;
    for (; n < kMessagesMax; n++) if (!r->messages[n]) break;
    if (n == kMessagesMax) {
        free(r->messages[0]);
        for (size_t i = 0; i < kMessagesMax - 1; i++) r->messages[i] = r->messages[i + 1];
        r->messages[kMessagesMax - 1] = nullptr;
        n = kMessagesMax - 1;
    }
    r->messages[n] = (char*)malloc(REB_STRING_SIZE_MAX);
    r->messages[n][0] = type;
    strncpy(r->messages[n] + 1, msg, REB_STRING_SIZE_MAX - 2);
    r->messages[n][REB_STRING_SIZE_MAX - 1] = '\0';
}

REB_API void reb_simulation_info(struct reb_simulation* const r, const char* const msg) { push_message(r, 'i', msg); }
REB_API void reb_simulation_warning(struct reb_simulation* const r, const char* const msg) { push_message(r, 'w', msg); }
REB_API void reb_simulation_error(struct reb_simulation* const r, const char* const msg) {
    set_last_error(msg);
    push_message(r, 'e', msg);
}

REB_API int reb_pop_message(char** messages, char* const buf) {
    if (messages && messages[0]) {
        char* w0 = messages[0];
        for (size_t i = 0; i < kMessagesMax - 1; i++) messages[i] = messages[i + 1];
        messages[kMessagesMax - 1] = nullptr;
        strcpy(buf, w0);
        free(w0);
        return 1;
    }
    return 0;
}

REB_API void reb_exit(const char* const msg) {
    push_message(nullptr, 'e', msg);
    exit(EXIT_FAILURE);
}

REB_API void reb_simulation_stop(struct reb_simulation* const r) { r->status = REB_STATUS_USER; }

// rebound.c:215-217: memory the library handed out (strings, streams).
REB_API void reb_free(void* p) { free(p); }

// ===================================================================== registry
// Names and behaviour follow rebound.c:49-113 and simulation.c:201-238.  The
// built-in names that are not part of this engine stay selectable so existing
// code keeps linking; stepping with them records an error.
void rb_leapfrog_step(struct reb_simulation* r, void* state);
void* rb_leapfrog_create(void);
void rb_sei_step(struct reb_simulation* r, void* state);
void* rb_sei_create(void);
void rb_whfast_step(struct reb_simulation* r, void* state);
void rb_whfast_synchronize(struct reb_simulation* r, void* state);
void* rb_whfast_create(void);
void rb_whfast_free(void* state);

static void rb_plain_free(void* p) { free(p); }
static void rb_none_step(struct reb_simulation* r, void* state) {
    (void)state;
    r->t += r->dt;
    r->dt_last_done = r->dt;
}
static void rb_unsupported_step(struct reb_simulation* r, void* state) {
    (void)state;
    char msg[REB_STRING_SIZE_MAX];
    snprintf(msg, sizeof(msg),
             "Integrator '%s' is not part of the MI355X force-and-step engine. Use \"leapfrog\", \"sei\" or \"whfast\".",
             r->integrator.name ? r->integrator.name : "(null)");
    reb_simulation_error(r, msg);
}

const char* reb_build_str = __DATE__ " " __TIME__;
const char* reb_version_str = "5.0.0";
const char* reb_githash_str = "rebound-amd";

const struct reb_integrator reb_integrator_none = {
    (char*)"Does nothing other than advance time.", rb_none_step, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
static const struct reb_integrator rb_integrator_leapfrog = {
    (char*)"Leapfrog (drift-kick-drift), orders 2/4/6/8, force evaluation on the GPU.",
    rb_leapfrog_step, nullptr, rb_leapfrog_create, rb_plain_free, nullptr, nullptr, nullptr};
static const struct reb_integrator rb_integrator_sei = {
    (char*)"Symplectic Epicycle Integrator for the shearing sheet, on the GPU.",
    rb_sei_step, nullptr, rb_sei_create, rb_plain_free, nullptr, nullptr, nullptr};
static const struct reb_integrator rb_integrator_whfast = {
    (char*)"WHFast (Jacobi or democratic heliocentric coordinates, symplectic correctors), on the GPU.",
    rb_whfast_step, rb_whfast_synchronize, rb_whfast_create, rb_whfast_free, nullptr, nullptr, nullptr};
static const struct reb_integrator rb_integrator_unsupported = {
    (char*)"Not available in this engine.", rb_unsupported_step, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};

struct BuiltinEntry { const char* name; const struct reb_integrator* cb; };
static const BuiltinEntry kBuiltins[] = {
    {"ias15", &rb_integrator_unsupported}, {"whfast", &rb_integrator_whfast}, {"sei", &rb_integrator_sei},
    {"leapfrog", &rb_integrator_leapfrog}, {"janus", &rb_integrator_unsupported},
    {"mercurius", &rb_integrator_unsupported}, {"saba", &rb_integrator_unsupported},
    {"eos", &rb_integrator_unsupported}, {"bs", &rb_integrator_unsupported},
    {"whfast512", &rb_integrator_unsupported}, {"trace", &rb_integrator_unsupported},
    {"none", &reb_integrator_none},
};
static const size_t kNumBuiltins = sizeof(kBuiltins) / sizeof(kBuiltins[0]);

static std::mutex g_registry_mutex;
static std::vector<struct reb_integrator_configuration> g_custom;

static void* set_integrator(struct reb_simulation* r, const char* name, struct reb_integrator cb) {
    if (r->integrator.callbacks.free && r->integrator.state) r->integrator.callbacks.free(r->integrator.state);
    r->integrator.callbacks = cb;
    r->integrator.name = name;
    r->integrator.state = cb.create ? cb.create() : nullptr;
    return r->integrator.state;
}

REB_API void* reb_simulation_set_integrator(struct reb_simulation* r, const char* name) {
    if (r->integrator.name && !r->is_synchronized)
        reb_simulation_warning(r, "Changing integrators while simulation is not synchronized results in undefined behaviour.");
    for (size_t i = 0; i < kNumBuiltins; i++)
        if (strcmp(name, kBuiltins[i].name) == 0) return set_integrator(r, kBuiltins[i].name, *kBuiltins[i].cb);
    {
        std::lock_guard<std::mutex> g(g_registry_mutex);
        for (auto& c : g_custom)
            if (strcmp(c.name, name) == 0) return set_integrator(r, c.name, c.callbacks);
    }
    reb_simulation_error(r, "Integrator not found.");
    return nullptr;
}

REB_API void reb_integrator_register(const struct reb_integrator integrator, const char* name) {
    if (name == nullptr || strnlen(name, 256) == 0 || strnlen(name, 256) == 256) {
        reb_simulation_error(nullptr, "Must provide integrator name when registering.");
        return;
    }
    for (size_t i = 0; i < kNumBuiltins; i++)
        if (strcmp(name, kBuiltins[i].name) == 0) {
            reb_simulation_error(nullptr, "Integrator name must be unique but name already exists.");
            return;
        }
    std::lock_guard<std::mutex> g(g_registry_mutex);
    for (auto& c : g_custom)
        if (strcmp(c.name, name) == 0) {
            reb_simulation_error(nullptr, "Integrator name must be unique but name already exists.");
            return;
        }
    struct reb_integrator_configuration c;
    c.name = strdup(name);
    c.callbacks = integrator;
    c.state = nullptr;
    g_custom.push_back(c);
}

REB_API char** reb_integrators_registered(void) {
    std::lock_guard<std::mutex> g(g_registry_mutex);
    char** names = (char**)malloc(sizeof(char*) * (kNumBuiltins + g_custom.size() + 1));
    size_t n = 0;
    for (size_t i = 0; i < kNumBuiltins; i++) names[n++] = (char*)kBuiltins[i].name;
    for (auto& c : g_custom) names[n++] = (char*)c.name;
    names[n] = nullptr;
    return names;
}

// ===================================================================== lifecycle
// simulation.c:98-122: defaults of a new simulation.
static unsigned int time_seed() {
    struct timeval tim;
    gettimeofday(&tim, nullptr);
    return (unsigned int)(tim.tv_usec + getpid());
}

REB_API size_t reb_simulation_struct_size(void) { return sizeof(struct reb_simulation); }

REB_API struct reb_simulation* reb_simulation_create(void) {
    SimBlock* b = (SimBlock*)calloc(1, sizeof(SimBlock));
    new (&b->dev) DeviceState();
    struct reb_simulation* r = &b->pub;
    r->rand_seed = time_seed();
    r->G = 1;
    r->dt = 0.001;
    r->gravity = REB_GRAVITY_BASIC;
    r->root_size = -1;
    r->opening_angle2 = 0.25;
    r->OMEGAZ = -1.0;
    r->N_root_x = 1;
    r->N_root_y = 1;
    r->N_root_z = 1;
    r->N_active = SIZE_MAX;
    r->N_targets = SIZE_MAX;
    r->exact_finish_time = 1;
    r->output_timing_last = -1;
    r->simulationarchive_version = 5;
    r->is_synchronized = 1;
    reb_simulation_set_integrator(r, "ias15");
    return r;
}

REB_API void reb_simulation_free(struct reb_simulation* const r) {
    if (!r) return;
    if (r->integrator.callbacks.free) r->integrator.callbacks.free(r->integrator.state);
    if (r->extras_cleanup) r->extras_cleanup(r);
    device_release(r);
    if (r->messages) {
        for (size_t i = 0; i < kMessagesMax; i++) free(r->messages[i]);
        free(r->messages);
    }
    if (r->free_particle_ap)
        for (size_t i = 0; i < r->N; i++) r->free_particle_ap(&r->particles[i]);
    free(r->particles);
    collisions_free(r);
    free(r->gravity_cs);
    free(r->simulationarchive_filename);
    SimBlock* b = reinterpret_cast<SimBlock*>(r);
    b->dev.~DeviceState();
    free(b);
}

// Copy of the host-visible state: particles, scalar settings and integrator
// options.  (The reference copies through its binary format; this engine
// copies the struct and deep-copies owned arrays.)
REB_API struct reb_simulation* reb_simulation_copy(struct reb_simulation* r) {
    ensure_host(r);
    struct reb_simulation* c = reb_simulation_create();
    struct reb_integrator_configuration keep = c->integrator;
    if (keep.callbacks.free && keep.state) keep.callbacks.free(keep.state);
    memcpy(c, r, sizeof(struct reb_simulation));
    c->messages = nullptr;
    c->particles = (struct reb_particle*)malloc(sizeof(struct reb_particle) * (r->N_allocated ? r->N_allocated : 1));
    if (r->N) memcpy(c->particles, r->particles, sizeof(struct reb_particle) * r->N);
    for (size_t i = 0; i < r->N; i++) c->particles[i].sim = c;
    c->collisions = nullptr;
    c->N_allocated_collisions = 0;
    c->N_collisions = 0;
    c->gravity_cs = nullptr;
    c->N_allocated_gravity_cs = 0;
    c->tree_root = nullptr;
    c->simulationarchive_filename = r->simulationarchive_filename ? strdup(r->simulationarchive_filename) : nullptr;
    c->integrator.state = nullptr;
    if (r->integrator.callbacks.create) {
        c->integrator.state = r->integrator.callbacks.create();
        if (strcmp(r->integrator.name, "leapfrog") == 0)
            memcpy(c->integrator.state, r->integrator.state, sizeof(struct reb_integrator_leapfrog_state));
        if (strcmp(r->integrator.name, "whfast") == 0) {
            auto* s = (struct reb_integrator_whfast_state*)r->integrator.state;
            auto* d = (struct reb_integrator_whfast_state*)c->integrator.state;
            d->corrector = s->corrector;
            d->corrector2 = s->corrector2;
            d->kernel = s->kernel;
            d->coordinates = s->coordinates;
            d->safe_mode = s->safe_mode;
            d->keep_unsynchronized = s->keep_unsynchronized;
        }
    }
    dev_of(c)->precision = dev_of(r)->precision;
    dev_of(c)->cor_constant = dev_of(r)->cor_constant;
    dev_of(c)->dev_forces = dev_of(r)->dev_forces;
    dev_of(c)->dev_forces_user = dev_of(r)->dev_forces_user;
    c->did_modify_particles = 1;
    return c;
}

// A constant coefficient of restitution as a C callback, so the host resolver
// (collision.c:573-665) does not call back into an interpreter per collision.
REB_API double reb_amd_coefficient_of_restitution_constant(const struct reb_simulation* const r, double v) {
    (void)v;
    return dev_of(r)->cor_constant;
}

REB_API void reb_amd_set_coefficient_of_restitution(struct reb_simulation* r, double eps) {
    dev_of(r)->cor_constant = eps;
    r->coefficient_of_restitution = reb_amd_coefficient_of_restitution_constant;
}

REB_API void reb_amd_set_device_forces(struct reb_simulation* r, reb_amd_device_forces_fn fn, void* user) {
    dev_of(r)->dev_forces = fn;
    dev_of(r)->dev_forces_user = user;
}

// ===================================================================== particles
// particle.c:47-76 (add, with the box check of boundary.c:203-238),
// particle.c:322-373 (remove).
static int particle_is_in_box(const struct reb_simulation* r, const struct reb_particle& p) {
    switch (r->boundary) {
        case REB_BOUNDARY_OPEN:
        case REB_BOUNDARY_SHEAR:
        case REB_BOUNDARY_PERIODIC: {
            const double bx = r->root_size * (double)r->N_root_x;
            const double by = r->root_size * (double)r->N_root_y;
            const double bz = r->root_size * (double)r->N_root_z;
            if (p.x > bx / 2. || p.x < -bx / 2.) return 0;
            if (p.y > by / 2. || p.y < -by / 2.) return 0;
            if (p.z > bz / 2. || p.z < -bz / 2.) return 0;
            return 1;
        }
        default:
            return 1;
    }
}

REB_API void reb_simulation_add(struct reb_simulation* const r, struct reb_particle pt) {
    if (particle_is_in_box(r, pt) == 0) {
        reb_simulation_error(r, "Particle outside of box boundaries. Did not add particle.");
        return;
    }
    ensure_host(r);
    if (r->N_allocated <= r->N) {
        size_t old = r->N_allocated;
        r->N_allocated = r->N_allocated ? r->N_allocated * 2 : 8;
        r->particles = (struct reb_particle*)realloc(r->particles, sizeof(struct reb_particle) * r->N_allocated);
        memset(r->particles + old, 0, (r->N_allocated - old) * sizeof(struct reb_particle));
    }
    r->particles[r->N] = pt;
    r->particles[r->N].sim = r;
    r->N++;
    if (r->integrator.callbacks.did_add_particle) r->integrator.callbacks.did_add_particle(r);
    r->did_modify_particles = 1;
    mark_host_newer(r);
}

REB_API void reb_simulation_remove_all_particles(struct reb_simulation* const r) {
    r->N = 0;
    r->N_allocated = 0;
    r->N_active = SIZE_MAX;
    r->N_var = 0;
    free(r->particles);
    r->particles = nullptr;
    free(r->particles_var);
    r->particles_var = nullptr;
    mark_host_newer(r);
}

REB_API int reb_simulation_remove_particle(struct reb_simulation* const r, size_t index) {
    if (r->N_var) {
        reb_simulation_error(r, "Removing particles not supported when variational particles are in use. Did not remove particle.");
        return 1;
    }
    ensure_host(r);
    if (r->integrator.callbacks.will_remove_particle) r->integrator.callbacks.will_remove_particle(r, index);
    if (r->N == 1) {
        r->N = 0;
        r->did_modify_particles = 1;
        if (r->free_particle_ap) r->free_particle_ap(&r->particles[index]);
        reb_simulation_warning(r, "Last particle removed.");
        mark_host_newer(r);
        return 0;
    }
    if (index >= r->N) {
        char w[REB_STRING_SIZE_MAX];
        snprintf(w, sizeof(w), "Index %zu passed to particles_remove was out of range (N=%zu).  Did not remove particle.", index, r->N);
        reb_simulation_error(r, w);
        return 1;
    }
    r->N--;
    if (r->free_particle_ap) r->free_particle_ap(&r->particles[index]);
    if (index < r->N_active && r->N_active != SIZE_MAX) r->N_active--;
    for (size_t j = index; j < r->N; j++) r->particles[j] = r->particles[j + 1];
    r->did_modify_particles = 1;
    mark_host_newer(r);
    return 0;
}

REB_API int reb_particle_cmp(struct reb_particle p1, struct reb_particle p2) {
    int differ = 0;
    differ = differ || (p1.x != p2.x) || (p1.y != p2.y) || (p1.z != p2.z);
    differ = differ || (p1.vx != p2.vx) || (p1.vy != p2.vy) || (p1.vz != p2.vz);
    differ = differ || (p1.ax != p2.ax) || (p1.ay != p2.ay) || (p1.az != p2.az);
    differ = differ || (p1.m != p2.m) || (p1.r != p2.r) || (p1.name != p2.name);
    return differ;
}

REB_API struct reb_particle reb_particle_nan(void) {
    struct reb_particle p;
    memset(&p, 0, sizeof(p));
    p.x = p.y = p.z = p.vx = p.vy = p.vz = p.ax = p.ay = p.az = p.m = p.r = nan("");
    return p;
}

// ===================================================================== random numbers
// glibc rand_r (three LCG steps combined into 31 bits; rebound.c:238-259 restates
// it) drawn from r->rand_seed, and the reference's transforms (tools.c:64-105).
static int rb_rand_r(unsigned int* seed) {
    unsigned int next = *seed;
    int result;
    next *= 1103515245u;
    next += 12345u;
    result = (unsigned int)(next / 65536) % 2048;
    next *= 1103515245u;
    next += 12345u;
    result <<= 10;
    result ^= (unsigned int)(next / 65536) % 1024;
    next *= 1103515245u;
    next += 12345u;
    result <<= 10;
    result ^= (unsigned int)(next / 65536) % 1024;
    *seed = next;
    return result;
}
static const double kRandMax = 2147483647.;

REB_API double reb_random_uniform(struct reb_simulation* r, double min, double max) {
    unsigned int seed;
    unsigned int* seedp;
    if (r) {
        seedp = &r->rand_seed;
    } else {
        seed = time_seed();
        seedp = &seed;
    }
    return ((double)rb_rand_r(seedp)) / kRandMax * (max - min) + min;
}

REB_API double reb_random_powerlaw(struct reb_simulation* r, double min, double max, double slope) {
    double y = reb_random_uniform(r, 0., 1.);
    if (slope == -1) return exp(y * log(max / min) + log(min));
    return pow((pow(max, slope + 1.) - pow(min, slope + 1.)) * y + pow(min, slope + 1.), 1. / (slope + 1.));
}

REB_API double reb_random_normal(struct reb_simulation* r, double variance) {
    double v1 = 0., v2 = 0., rsq = 1.;
    unsigned int seed;
    unsigned int* seedp;
    if (r) {
        seedp = &r->rand_seed;
    } else {
        seed = time_seed();
        seedp = &seed;
    }
    while (rsq >= 1. || rsq < 1.0e-12) {
        v1 = 2. * ((double)rb_rand_r(seedp)) / kRandMax - 1.0;
        v2 = 2. * ((double)rb_rand_r(seedp)) / kRandMax - 1.0;
        rsq = v1 * v1 + v2 * v2;
    }
    return v1 * sqrt(-2. * log(rsq) / rsq * variance);
}

REB_API double reb_random_rayleigh(struct reb_simulation* r, double sigma) {
    double y = reb_random_uniform(r, 0., 1.);
    return sigma * sqrt(-2 * log(y));
}

// ===================================================================== generators
// tools.c:463-502: Plummer sphere (Aarseth, Henon & Wielen 1974).
REB_API void reb_simulation_add_plummer(struct reb_simulation* r, size_t _N, double M, double R) {
    double E = 3. / 64. * M_PI * M * M / R;
    for (size_t i = 0; i < _N; i++) {
        struct reb_particle star;
        memset(&star, 0, sizeof(star));
        double _r = pow(pow(reb_random_uniform(r, 0, 1), -2. / 3.) - 1., -1. / 2.);
        double x2 = reb_random_uniform(r, 0, 1);
        double x3 = reb_random_uniform(r, 0, 2. * M_PI);
        star.z = (1. - 2. * x2) * _r;
        star.x = sqrt(_r * _r - star.z * star.z) * cos(x3);
        star.y = sqrt(_r * _r - star.z * star.z) * sin(x3);
        double x5, g, q;
        do {
            x5 = reb_random_uniform(r, 0., 1.);
            q = reb_random_uniform(r, 0., 1.);
            g = q * q * pow(1. - q * q, 7. / 2.);
        } while (0.1 * x5 > g);
        double ve = pow(2., 1. / 2.) * pow(1. + _r * _r, -1. / 4.);
        double v = q * ve;
        double x6 = reb_random_uniform(r, 0., 1.);
        double x7 = reb_random_uniform(r, 0., 2. * M_PI);
        star.vz = (1. - 2. * x6) * v;
        star.vx = sqrt(v * v - star.vz * star.vz) * cos(x7);
        star.vy = sqrt(v * v - star.vz * star.vz) * sin(x7);
        star.x *= 3. * M_PI / 64. * M * M / E;
        star.y *= 3. * M_PI / 64. * M * M / E;
        star.z *= 3. * M_PI / 64. * M * M / E;
        star.vx *= sqrt(E * 64. / 3. / M_PI / M);
        star.vy *= sqrt(E * 64. / 3. / M_PI / M);
        star.vz *= sqrt(E * 64. / 3. / M_PI / M);
        star.m = M / (double)_N;
        reb_simulation_add(r, star);
    }
}

// tools.c:940-1002: Cartesian state from orbital elements (Murray & Dermott 2.122, 2.36).
REB_API struct reb_particle reb_particle_from_orbit_err(double G, struct reb_particle primary, double m, double a,
                                                        double e, double inc, double Omega, double omega, double f,
                                                        int* err) {
    if (e == 1.) { *err = 1; return reb_particle_nan(); }
    if (e < 0.) { *err = 2; return reb_particle_nan(); }
    if (e > 1.) {
        if (a > 0.) { *err = 3; return reb_particle_nan(); }
    } else {
        if (a < 0.) { *err = 4; return reb_particle_nan(); }
    }
    // The reference library, built by gcc, evaluates each angle's sine and
    // cosine with one glibc sincos() call; the same calls give the same bits.
    double sf, cf;
    sincos(f, &sf, &cf);
    if (e * cf < -1.) { *err = 5; return reb_particle_nan(); }
    if (primary.m < 1.E-308) { *err = 6; return reb_particle_nan(); }

    struct reb_particle p;
    memset(&p, 0, sizeof(p));
    p.m = m;
    double r = a * (1 - e * e) / (1 + e * cf);
    double v0 = sqrt(G * (m + primary.m) / a / (1. - e * e));
    double sO, cO, so, co, si, ci;
    sincos(Omega, &sO, &cO);
    sincos(omega, &so, &co);
    sincos(inc, &si, &ci);
    p.x = primary.x + r * (cO * (co * cf - so * sf) - sO * (so * cf + co * sf) * ci);
    p.y = primary.y + r * (sO * (co * cf - so * sf) + cO * (so * cf + co * sf) * ci);
    p.z = primary.z + r * (so * cf + co * sf) * si;
    p.vx = primary.vx + v0 * ((e + cf) * (-ci * co * sO - cO * so) - sf * (co * cO - ci * so * sO));
    p.vy = primary.vy + v0 * ((e + cf) * (ci * co * cO - sO * so) - sf * (co * sO + ci * so * cO));
    p.vz = primary.vz + v0 * ((e + cf) * co * si - sf * si * so);
    p.ax = 0;
    p.ay = 0;
    p.az = 0;
    return p;
}

REB_API struct reb_particle reb_particle_from_orbit(double G, struct reb_particle primary, double m, double a, double e,
                                                    double i, double Omega, double omega, double f) {
    int err;
    return reb_particle_from_orbit_err(G, primary, m, a, e, i, Omega, omega, f, &err);
}

// ===================================================================== centre of mass
// tools.c:370-420.  Sequential pairwise mass-weighted merge, as the reference.
REB_API struct reb_particle reb_particle_com_of_pair(struct reb_particle p1, struct reb_particle p2) {
    p1.x = p1.x * p1.m + p2.x * p2.m;
    p1.y = p1.y * p1.m + p2.y * p2.m;
    p1.z = p1.z * p1.m + p2.z * p2.m;
    p1.vx = p1.vx * p1.m + p2.vx * p2.m;
    p1.vy = p1.vy * p1.m + p2.vy * p2.m;
    p1.vz = p1.vz * p1.m + p2.vz * p2.m;
    p1.ax = p1.ax * p1.m + p2.ax * p2.m;
    p1.ay = p1.ay * p1.m + p2.ay * p2.m;
    p1.az = p1.az * p1.m + p2.az * p2.m;
    p1.m += p2.m;
    if (p1.m > 0.) {
        p1.x /= p1.m;
        p1.y /= p1.m;
        p1.z /= p1.m;
        p1.vx /= p1.m;
        p1.vy /= p1.m;
        p1.vz /= p1.m;
        p1.ax /= p1.m;
        p1.ay /= p1.m;
        p1.az /= p1.m;
    }
    return p1;
}

REB_API struct reb_particle reb_simulation_com_range(struct reb_simulation* r, size_t first, size_t last) {
    ensure_host(r);
    struct reb_particle com;
    memset(&com, 0, sizeof(com));
    for (size_t i = first; i < last; i++) com = reb_particle_com_of_pair(com, r->particles[i]);
    return com;
}

REB_API struct reb_particle reb_simulation_com(struct reb_simulation* r) { return reb_simulation_com_range(r, 0, r->N); }

REB_API void reb_simulation_move_to_com(struct reb_simulation* const r) {
    if (r->N_var_config) {
        reb_simulation_error(r, "Variational particles are not supported by this engine.");
        return;
    }
    struct reb_particle com = reb_simulation_com(r);
    for (size_t i = 0; i < r->N; i++) {
        r->particles[i].x -= com.x;
        r->particles[i].y -= com.y;
        r->particles[i].z -= com.z;
        r->particles[i].vx -= com.vx;
        r->particles[i].vy -= com.vy;
        r->particles[i].vz -= com.vz;
    }
    mark_host_newer(r);
    // tools.c:365: wrap into the box afterwards
    if (r->boundary != REB_BOUNDARY_NONE && r->N) {
        if (ensure_device(r) && boundary_device(r)) ensure_host(r);
    }
}

REB_API void reb_simulation_move_to_hel(struct reb_simulation* const r) {
    ensure_host(r);
    if (r->N > 0) {
        struct reb_particle hel = r->particles[0];
        for (size_t i = 1; i < r->N; i++) {
            r->particles[i].x -= hel.x;
            r->particles[i].y -= hel.y;
            r->particles[i].z -= hel.z;
            r->particles[i].vx -= hel.vx;
            r->particles[i].vy -= hel.vy;
            r->particles[i].vz -= hel.vz;
        }
        r->particles[0].x = r->particles[0].y = r->particles[0].z = 0.;
        r->particles[0].vx = r->particles[0].vy = r->particles[0].vz = 0.;
    }
    mark_host_newer(r);
}

REB_API struct reb_vec3d reb_simulation_angular_momentum(const struct reb_simulation* const r) {
    ensure_host((struct reb_simulation*)r);
    struct reb_vec3d L = {0, 0, 0};
    for (size_t i = 0; i < r->N; i++) {
        struct reb_particle pi = r->particles[i];
        L.x += pi.m * (pi.y * pi.vz - pi.z * pi.vy);
        L.y += pi.m * (pi.z * pi.vx - pi.x * pi.vz);
        L.z += pi.m * (pi.x * pi.vy - pi.y * pi.vx);
    }
    return L;
}

// ===================================================================== ghost boxes
// boundary.c:145-201.  Offsets evaluated on the host (libm fmod) once per
// force evaluation and passed to the kernels.
REB_API struct reb_vec6d reb_boundary_get_ghostbox(struct reb_simulation* const r, int i, int j, int k) {
    const double bx = r->root_size * (double)r->N_root_x;
    const double by = r->root_size * (double)r->N_root_y;
    const double bz = r->root_size * (double)r->N_root_z;
    struct reb_vec6d gb = {0, 0, 0, 0, 0, 0};
    switch (r->boundary) {
        case REB_BOUNDARY_OPEN:
        case REB_BOUNDARY_PERIODIC:
            gb.x = bx * (double)i;
            gb.y = by * (double)j;
            gb.z = bz * (double)k;
            return gb;
        case REB_BOUNDARY_SHEAR: {
            const double OMEGA = r->OMEGA;
            gb.vx = 0.;
            gb.vy = -1.5 * (double)i * OMEGA * bx;
            gb.vz = 0.;
            double shift;
            if (i == 0) {
                shift = -fmod(gb.vy * r->t, by);
            } else if (i > 0) {
                shift = -fmod(gb.vy * r->t - by / 2., by) - by / 2.;
            } else {
                shift = -fmod(gb.vy * r->t + by / 2., by) + by / 2.;
            }
            gb.x = bx * (double)i;
            gb.y = by * (double)j - shift;
            gb.z = bz * (double)k;
            return gb;
        }
        default:
            return gb;
    }
}

}  // extern "C"

namespace rbamd {
std::vector<struct reb_vec6d> ghostboxes(struct reb_simulation* r, int nx, int ny, int nz) {
    std::vector<struct reb_vec6d> v;
    for (int gbx = -nx; gbx <= nx; gbx++)
        for (int gby = -ny; gby <= ny; gby++)
            for (int gbz = -nz; gbz <= nz; gbz++) v.push_back(reb_boundary_get_ghostbox(r, gbx, gby, gbz));
    return v;
}
}  // namespace rbamd
