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
    free(r->particles_var);
    free(r->var_config);
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
    c->particles_var = nullptr;
    c->var_config = nullptr;
    if (r->N_var) {
        c->particles_var = (struct reb_particle*)malloc(sizeof(struct reb_particle) * r->N_var);
        memcpy(c->particles_var, r->particles_var, sizeof(struct reb_particle) * r->N_var);
        for (size_t i = 0; i < r->N_var; i++) c->particles_var[i].sim = c;
    }
    if (r->N_var_config) {
        c->var_config = (struct reb_variational_configuration*)malloc(sizeof(struct reb_variational_configuration) *
                                                                       r->N_var_config);
        memcpy(c->var_config, r->var_config, sizeof(struct reb_variational_configuration) * r->N_var_config);
        for (size_t v = 0; v < r->N_var_config; v++) c->var_config[v].sim = c;
    }
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

// tools.c:201-375: the variational sets are shifted by the variation of the
// centre of mass (second order first, then first order), then the bodies.
REB_API void reb_simulation_move_to_com(struct reb_simulation* const r) {
    struct reb_particle com = reb_simulation_com(r);
    const size_t N = r->N;
    struct reb_particle* const p = r->particles;
    struct reb_particle* const pv = r->particles_var;
    // component k of a particle: x y z vx vy vz
    auto c = [](const struct reb_particle& q, int k) {
        switch (k) {
            case 0: return q.x;
            case 1: return q.y;
            case 2: return q.z;
            case 3: return q.vx;
            case 4: return q.vy;
            default: return q.vz;
        }
    };
    auto shift = [&](size_t index, const double* s) {
        for (size_t i = 0; i < N; i++) {
            pv[i + index].x -= s[0];
            pv[i + index].y -= s[1];
            pv[i + index].z -= s[2];
            pv[i + index].vx -= s[3];
            pv[i + index].vy -= s[4];
            pv[i + index].vz -= s[5];
        }
    };
    // a set must lie inside particles_var (a file could say otherwise)
    auto fits = [&](long long i) { return i >= 0 && (size_t)i + N <= r->N_var; };
    for (size_t v = 0; v < r->N_var_config; v++) {
        const struct reb_variational_configuration& vc = r->var_config[v];
        if (vc.testparticle >= 0 || vc.order != 2) continue;
        if (!fits(vc.index) || !fits(vc.index_1st_order_a) || !fits(vc.index_1st_order_b)) continue;
        const size_t index = (size_t)vc.index, a = (size_t)vc.index_1st_order_a, b = (size_t)vc.index_1st_order_b;
        double dma = 0., dmb = 0., ddm = 0.;
        for (size_t i = 0; i < N; i++) {
            dma += pv[i + a].m;
            dmb += pv[i + b].m;
            ddm += pv[i + index].m;
        }
        double s[6] = {0., 0., 0., 0., 0., 0.};
        for (size_t i = 0; i < N; i++)
            for (int k = 0; k < 6; k++) {
                s[k] += c(pv[i + index], k) / com.m * p[i].m;
                s[k] += c(pv[i + a], k) / com.m * pv[i + b].m;
                s[k] -= c(pv[i + a], k) * p[i].m / com.m / com.m * dmb;
                s[k] += c(pv[i + b], k) / com.m * pv[i + a].m;
                s[k] += c(p[i], k) / com.m * pv[i + index].m;
                s[k] -= c(p[i], k) * pv[i + a].m / com.m / com.m * dmb;
                s[k] -= c(pv[i + b], k) * p[i].m / com.m / com.m * dma;
                s[k] -= c(p[i], k) * pv[i + b].m / com.m / com.m * dma;
                s[k] += 2. * c(p[i], k) * p[i].m / com.m / com.m / com.m * dma * dmb;
                s[k] -= c(p[i], k) * p[i].m / com.m / com.m * ddm;
            }
        shift(index, s);
    }
    for (size_t v = 0; v < r->N_var_config; v++) {
        const struct reb_variational_configuration& vc = r->var_config[v];
        if (vc.testparticle >= 0 || vc.order != 1 || !fits(vc.index)) continue;
        const size_t index = (size_t)vc.index;
        // tools.c:319-321 sums particles[i+index].m (the bodies' array, at
        // the set's offset); entries past the bodies count as zero here
        double dm = 0.;
        for (size_t i = 0; i < N; i++) dm += i + index < N ? p[i + index].m : 0.;
        double s[6] = {0., 0., 0., 0., 0., 0.};
        for (size_t i = 0; i < N; i++)
            for (int k = 0; k < 6; k++) {
                s[k] += p[i].m / com.m * c(pv[i + index], k);
                s[k] += c(p[i], k) / com.m * pv[i + index].m;
                s[k] -= c(p[i], k) / (com.m * com.m) * p[i].m * dm;
            }
        shift(index, s);
    }
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

// ---------------------------------------------------------------- variational sets
// tools.c:1382-1430
static int add_var_config(struct reb_simulation* r, int order, int testparticle, size_t a, size_t b) {
    r->N_var_config++;
    r->var_config = (struct reb_variational_configuration*)realloc(
        r->var_config, sizeof(struct reb_variational_configuration) * r->N_var_config);
    struct reb_variational_configuration& vc = r->var_config[r->N_var_config - 1];
    memset(&vc, 0, sizeof(vc));
    const size_t index = r->N_var;
    vc.sim = r;
    vc.order = order;
    vc.index = (int)index;
    vc.lrescale = 0;
    vc.testparticle = testparticle;
    vc.index_1st_order_a = (int)a;
    vc.index_1st_order_b = (int)b;
    const size_t n = testparticle >= 0 ? 1 : r->N;
    r->particles_var = (struct reb_particle*)realloc(r->particles_var, sizeof(struct reb_particle) * (r->N_var + n));
    for (size_t i = 0; i < n; i++) {
        memset(&r->particles_var[r->N_var + i], 0, sizeof(struct reb_particle));
        r->particles_var[r->N_var + i].sim = r;
    }
    r->N_var += n;
    return (int)index;
}

REB_API int reb_simulation_add_variation_1st_order(struct reb_simulation* const r, int testparticle) {
    return add_var_config(r, 1, testparticle, 0, 0);
}

REB_API int reb_simulation_add_variation_2nd_order(struct reb_simulation* const r, int testparticle,
                                                   size_t index_1st_order_a, size_t index_1st_order_b) {
    return add_var_config(r, 2, testparticle, index_1st_order_a, index_1st_order_b);
}

// tools.c:1313-1380, on the host copy of the variational particles (which
// the WHFast plugin keeps current after every step).
REB_API void reb_simulation_rescale_var(struct reb_simulation* const r) {
    if (r->N_var_config == 0) return;
    for (size_t v = 0; v < r->N_var_config; v++) {
        struct reb_variational_configuration* vc = &r->var_config[v];
        if (vc->lrescale < 0) continue;
        const size_t N = vc->testparticle < 0 ? r->N : 1;
        if (vc->index < 0 || (size_t)vc->index + N > r->N_var) continue;   // not inside particles_var
        double scale = 0;
        struct reb_particle* const q = r->particles_var + vc->index;
        for (size_t i = 0; i < N; i++) {
            const double f[6] = {fabs(q[i].x), fabs(q[i].y), fabs(q[i].z), fabs(q[i].vx), fabs(q[i].vy), fabs(q[i].vz)};
            for (double a : f) scale = a > scale ? a : scale;
        }
        if (!(scale > 1e100)) continue;
        if (vc->order == 1) {
            for (size_t w = 0; w < r->N_var_config; w++) {
                const struct reb_variational_configuration* wc = &r->var_config[w];
                if (wc->order == 2 && (wc->index_1st_order_a == vc->index || wc->index_1st_order_b == vc->index) &&
                    !(r->messages_var_rescale_warning & 4)) {
                    r->messages_var_rescale_warning |= 4;
                    reb_simulation_warning(r, "Rescaling a set of variational equations of order 1 which are being used by a set of variational equations of order 2. Order 2 equations will no longer be valid.");
                }
            }
        } else {
            if (!(r->messages_var_rescale_warning & 2)) {
                r->messages_var_rescale_warning |= 2;
                reb_simulation_warning(r, "Variational particles which are part of a second order variational equation have now large coordinates which might exceed range of floating point number range. REBOUND cannot rescale a second order variational equation as it is non-linear.");
            }
            return;
        }
        if (!r->is_synchronized) {
            if (!(r->messages_var_rescale_warning & 1)) {
                r->messages_var_rescale_warning |= 1;
                reb_simulation_warning(r, "Variational particles have large coordinates which might exceed range of floating point numbers. Rescaling failed because integrator was not synchronized. Turn on safe_mode or manually synchronize. Then rescale.");
            }
            return;
        }
        vc->lrescale += log(scale);
        for (size_t i = 0; i < N; i++) {
            q[i].x /= scale;
            q[i].y /= scale;
            q[i].z /= scale;
            q[i].vx /= scale;
            q[i].vy /= scale;
            q[i].vz /= scale;
        }
        r->did_modify_particles = 1;
    }
}

// tools.c:1431-1470
REB_API void reb_simulation_init_megno(struct reb_simulation* const r) {
    r->megno_Ys = 0.;
    r->megno_Yss = 0.;
    r->megno_cov_Yt = 0.;
    r->megno_var_t = 0.;
    r->megno_n = 0;
    r->megno_mean_Y = 0;
    r->megno_initial_t = r->t;
    r->megno_mean_t = 0;
    reb_simulation_add_variation_1st_order(r, -1);
    r->calculate_megno = 1;
    struct reb_particle* const q = r->particles_var;
    for (size_t i = 0; i < r->N; i++) {
        q[i].m = 0.;
        q[i].x = reb_random_normal(r, 1.);
        q[i].y = reb_random_normal(r, 1.);
        q[i].z = reb_random_normal(r, 1.);
        q[i].vx = reb_random_normal(r, 1.);
        q[i].vy = reb_random_normal(r, 1.);
        q[i].vz = reb_random_normal(r, 1.);
        const double deltad = 1. / sqrt(q[i].x * q[i].x + q[i].y * q[i].y + q[i].z * q[i].z + q[i].vx * q[i].vx +
                                        q[i].vy * q[i].vy + q[i].vz * q[i].vz);
        q[i].x *= deltad;
        q[i].y *= deltad;
        q[i].z *= deltad;
        q[i].vx *= deltad;
        q[i].vy *= deltad;
        q[i].vz *= deltad;
    }
}

REB_API void reb_simulation_init_megno_seed(struct reb_simulation* const r, unsigned int seed) {
    r->rand_seed = seed;
    reb_simulation_init_megno(r);
}

REB_API double reb_simulation_megno(struct reb_simulation* r) {
    if (r->t == r->megno_initial_t) return 0.;
    return r->megno_Yss / (r->t - r->megno_initial_t);
}

REB_API double reb_simulation_lyapunov(struct reb_simulation* r) {
    if (r->megno_var_t == 0.0) return 0.;
    return r->megno_cov_Yt / r->megno_var_t;
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
// tools.c:1504-1522: the running MEGNO sums after a step (dY from the device)
void megno_update(struct reb_simulation* r, double dY, double dt_done) {
    r->megno_Ys += dY;
    const double Y = r->megno_Ys / (r->t - r->megno_initial_t);
    r->megno_Yss += Y * dt_done;
    r->megno_n++;
    const double _d_t = r->t - r->megno_initial_t - r->megno_mean_t;
    r->megno_mean_t += _d_t / (double)r->megno_n;
    const double _d_Y = reb_simulation_megno(r) - r->megno_mean_Y;
    r->megno_mean_Y += _d_Y / (double)r->megno_n;
    r->megno_cov_Yt += ((double)r->megno_n - 1.) / (double)r->megno_n * (r->t - r->megno_initial_t - r->megno_mean_t) *
                       (reb_simulation_megno(r) - r->megno_mean_Y);
    r->megno_var_t += ((double)r->megno_n - 1.) / (double)r->megno_n * (r->t - r->megno_initial_t - r->megno_mean_t) *
                      (r->t - r->megno_initial_t - r->megno_mean_t);
}
}  // namespace rbamd


namespace rbamd {
std::vector<struct reb_vec6d> ghostboxes(struct reb_simulation* r, int nx, int ny, int nz) {
    std::vector<struct reb_vec6d> v;
    for (int gbx = -nx; gbx <= nx; gbx++)
        for (int gby = -ny; gby <= ny; gby++)
            for (int gbz = -nz; gbz <= nz; gbz++) v.push_back(reb_boundary_get_ghostbox(r, gbx, gby, gbz));
    return v;
}
}  // namespace rbamd
