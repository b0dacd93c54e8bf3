// Binary files and simulation archives in the reference's format
// (binarydata.c, simulationarchive.c), so that checkpoints move between this
// engine and REBOUND in both directions.
//
// A file is a 64-byte header ("REBOUND Binary File. Version: <v>", NUL, git
// hash), then fields { uint64 size_name; uint64 size_data; name; data } in
// the order of the field tables below, a "functionpointers" flag, an "end"
// field and a 12-byte blob { int32 index, offset_prev, offset_next }.  Every
// further snapshot appended to an archive is the diff of the simulation
// against the first snapshot (the fields that changed or are new), an "end"
// field and a blob; snapshot k is the first snapshot with diff k applied.
//
// The tables list the reference's field names, types and order
// (binarydata.c:42-115 for the simulation; integrator_leapfrog.c:50-54,
// integrator_sei.c:62-69, integrator_whfast.c:71-99 for the integrators); the
// offsets are those of this engine's identical struct layouts.
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "internal.h"

// simulationarchive.h:30-43: the archive index, in the reference's layout
// (its Python package mirrors the struct field by field).
struct reb_simulationarchive {
    FILE* inf;                 // the archive's stream, kept open
    char* filename;            // NULL for an archive read from memory
    int version;
    int reb_version_major;
    int reb_version_minor;
    int reb_version_patch;
    double auto_interval;
    double auto_walltime;
    uint64_t auto_step;
    int64_t nblobs;            // snapshots, the first one included
    uint64_t* offset;          // byte offset of each snapshot
    double* t;                 // simulation time of each snapshot
};

namespace rbamd {

void whfast_refresh_host(struct reb_simulation* r);   // whfast.hip

namespace {

enum class T { DOUBLE, INT, UINT, UINT32, INT64, UINT64, SIZE_T, STRING, POINTER, CHARP_LIST };

struct Desc {
    const char* name;
    T type;
    size_t offset;
    size_t offset_n;     // POINTER / CHARP_LIST: offset of the element count (SIZE_MAX: one element if set)
    size_t element;      // POINTER: bytes per element
};

struct BinField {
    uint64_t size_name;
    uint64_t size_data;
};

struct Blob {
    int32_t index;
    int32_t offset_prev;
    int32_t offset_next;
};

constexpr uint64_t kHeaderTag = 0x20444E554F424552ull;   // "REBOUND " read as a field's size_name
constexpr size_t kNoCount = SIZE_MAX;

#define SIMF(n, t) Desc{#n, T::t, offsetof(struct reb_simulation, n), 0, 0}
const Desc kSimFields[] = {
    SIMF(t, DOUBLE),
    SIMF(G, DOUBLE),
    SIMF(softening, DOUBLE),
    SIMF(dt, DOUBLE),
    SIMF(N, SIZE_T),
    SIMF(N_var, SIZE_T),
    SIMF(N_active, SIZE_T),
    SIMF(testparticle_type, INT),
    SIMF(opening_angle2, DOUBLE),
    SIMF(status, INT),
    SIMF(exact_finish_time, INT),
    SIMF(force_is_velocity_dependent, UINT),
    SIMF(gravity_ignore_terms, UINT),
    SIMF(output_timing_last, DOUBLE),
    SIMF(save_messages, INT),
    SIMF(exit_max_distance, DOUBLE),
    SIMF(exit_min_distance, DOUBLE),
    SIMF(usleep, DOUBLE),
    SIMF(track_energy_offset, INT),
    SIMF(energy_offset, DOUBLE),
    SIMF(root_size, DOUBLE),
    SIMF(N_root_x, SIZE_T),
    SIMF(N_root_y, SIZE_T),
    SIMF(N_root_z, SIZE_T),
    SIMF(N_ghost_x, INT),
    SIMF(N_ghost_y, INT),
    SIMF(N_ghost_z, INT),
    SIMF(minimum_collision_velocity, DOUBLE),
    SIMF(collisions_plog, DOUBLE),
    SIMF(collisions_log_n, INT64),
    SIMF(calculate_megno, INT),
    SIMF(megno_Ys, DOUBLE),
    SIMF(megno_Yss, DOUBLE),
    SIMF(megno_cov_Yt, DOUBLE),
    SIMF(megno_var_t, DOUBLE),
    SIMF(megno_mean_t, DOUBLE),
    SIMF(megno_mean_Y, DOUBLE),
    SIMF(megno_initial_t, DOUBLE),
    SIMF(megno_n, INT64),
    SIMF(simulationarchive_auto_interval, DOUBLE),
    SIMF(simulationarchive_auto_walltime, DOUBLE),
    SIMF(simulationarchive_next, DOUBLE),
    SIMF(collision, INT),
    Desc{"integrator.name", T::STRING, offsetof(struct reb_simulation, integrator.name), 0, 0},
    SIMF(boundary, INT),
    SIMF(gravity, INT),
    SIMF(OMEGA, DOUBLE),
    SIMF(OMEGAZ, DOUBLE),
    SIMF(is_synchronized, UINT),
    SIMF(did_modify_particles, UINT),
    Desc{"particles", T::POINTER, offsetof(struct reb_simulation, particles), offsetof(struct reb_simulation, N),
         sizeof(struct reb_particle)},
    // variational particles and their sets (binarydata.c:94-95); the sets
    // are stored as their struct, as the reference does
    Desc{"particles_var", T::POINTER, offsetof(struct reb_simulation, particles_var),
         offsetof(struct reb_simulation, N_var), sizeof(struct reb_particle)},
    Desc{"var_config", T::POINTER, offsetof(struct reb_simulation, var_config),
         offsetof(struct reb_simulation, N_var_config), sizeof(struct reb_variational_configuration)},
    SIMF(simulationarchive_version, INT),
    SIMF(walltime, DOUBLE),
    SIMF(walltime_last_steps, DOUBLE),
    SIMF(python_unit_l, UINT32),
    SIMF(python_unit_m, UINT32),
    SIMF(python_unit_t, UINT32),
    SIMF(simulationarchive_auto_step, UINT64),
    SIMF(simulationarchive_next_step, UINT64),
    SIMF(steps_done, UINT64),
    SIMF(dt_last_done, DOUBLE),
    SIMF(rand_seed, UINT),
    SIMF(testparticle_hidewarnings, INT),
    Desc{"name_list", T::CHARP_LIST, offsetof(struct reb_simulation, name_list),
         offsetof(struct reb_simulation, N_name_list), 0},
};
#undef SIMF

struct SeiState {
    double lastdt, sindt, tandt, sindtz, tandtz;
};

const Desc kLeapfrogFields[] = {
    Desc{"order", T::UINT, offsetof(struct reb_integrator_leapfrog_state, order), 0, 0},
};
const Desc kSeiFields[] = {
    Desc{"lastdt", T::DOUBLE, offsetof(SeiState, lastdt), 0, 0},
    Desc{"sindt", T::DOUBLE, offsetof(SeiState, sindt), 0, 0},
    Desc{"tandt", T::DOUBLE, offsetof(SeiState, tandt), 0, 0},
    Desc{"sindtz", T::DOUBLE, offsetof(SeiState, sindtz), 0, 0},
    Desc{"tandtz", T::DOUBLE, offsetof(SeiState, tandtz), 0, 0},
};
#define WHF(n, t) Desc{#n, T::t, offsetof(struct reb_integrator_whfast_state, n), 0, 0}
const Desc kWhfastFields[] = {
    WHF(corrector, UINT),
    WHF(safe_mode, UINT),
    WHF(coordinates, INT),
    WHF(corrector2, UINT),
    WHF(kernel, INT),
    WHF(keep_unsynchronized, UINT),
    Desc{"p_jh", T::POINTER, offsetof(struct reb_integrator_whfast_state, p_jh),
         offsetof(struct reb_integrator_whfast_state, N_allocated), sizeof(struct reb_particle)},
    Desc{"p_jh_var", T::POINTER, offsetof(struct reb_integrator_whfast_state, p_jh_var),
         offsetof(struct reb_integrator_whfast_state, N_allocated_var), sizeof(struct reb_particle)},
};
#undef WHF

template <size_t K>
std::vector<Desc> list_of(const Desc (&a)[K]) {
    return std::vector<Desc>(a, a + K);
}

// The integrator's field table (the built-ins of this engine).
std::vector<Desc> integrator_fields(const char* name) {
    if (!name) return {};
    if (strcmp(name, "leapfrog") == 0) return list_of(kLeapfrogFields);
    if (strcmp(name, "sei") == 0) return list_of(kSeiFields);
    if (strcmp(name, "whfast") == 0) return list_of(kWhfastFields);
    return {};
}

size_t simple_size(T t) {
    switch (t) {
        case T::DOUBLE: return sizeof(double);
        case T::INT: return sizeof(int);
        case T::UINT: return sizeof(unsigned int);
        case T::UINT32: return sizeof(uint32_t);
        case T::INT64: return sizeof(int64_t);
        case T::UINT64: return sizeof(uint64_t);
        case T::SIZE_T: return sizeof(size_t);
        default: return 0;
    }
}

void put(std::string& buf, const void* p, size_t n) { buf.append(static_cast<const char*>(p), n); }

void put_field(std::string& buf, const std::string& name, const void* data, size_t n) {
    const BinField f{name.size() + 1, n};
    put(buf, &f, sizeof(f));
    put(buf, name.c_str(), name.size() + 1);
    if (n) put(buf, data, n);
}

// binarydata.c:490-623
void output_fields(std::string& buf, const std::vector<Desc>& list, const char* base, const std::string& prefix) {
    for (const Desc& d : list) {
        const std::string name = prefix.empty() ? std::string(d.name) : prefix + "." + d.name;
        const char* at = base + d.offset;
        switch (d.type) {
            case T::STRING: {
                const char* s = *reinterpret_cast<const char* const*>(at);
                put_field(buf, name, s, s ? strlen(s) + 1 : 0);
                break;
            }
            case T::POINTER: {
                const void* p = *reinterpret_cast<const void* const*>(at);
                size_t n = 0;
                if (d.offset_n != kNoCount)
                    n = *reinterpret_cast<const size_t*>(base + d.offset_n);
                else
                    n = p ? 1 : 0;
                const size_t bytes = n * d.element;
                if (bytes && p) put_field(buf, name, p, bytes);
                break;
            }
            case T::CHARP_LIST: {
                const size_t n = *reinterpret_cast<const size_t*>(base + d.offset_n);
                char* const* list = *reinterpret_cast<char* const* const*>(at);
                std::string data;
                for (size_t i = 0; i < n; i++) {
                    put(data, list[i], strlen(list[i]) + 1);
                    put(data, &list[i], sizeof(char*));   // the original pointer, matched on input
                }
                if (!data.empty()) put_field(buf, name, data.data(), data.size());
                break;
            }
            default:
                put_field(buf, name, at, simple_size(d.type));
                break;
        }
    }
}

// binarydata.c:625-678: a whole snapshot, header to blob.
std::string serialize(struct reb_simulation* r) {
    std::string buf;
    char header[64] = {0};
    const int n = snprintf(header, sizeof(header), "REBOUND Binary File. Version: %s", reb_version_str);
    snprintf(header + n + 1, sizeof(header) - n - 1, "%s", reb_githash_str);
    put(buf, header, 64);
    output_fields(buf, list_of(kSimFields), reinterpret_cast<const char*>(r), "");
    if (r->integrator.name && r->integrator.name[0] && r->integrator.state)
        output_fields(buf, integrator_fields(r->integrator.name), static_cast<const char*>(r->integrator.state),
                      std::string("integrator.") + r->integrator.name);
    const int fp = (r->coefficient_of_restitution || r->collision_resolve || r->additional_forces || r->heartbeat ||
                    r->post_timestep_modifications || r->free_particle_ap)
                       ? 1
                       : 0;
    put_field(buf, "functionpointers", &fp, sizeof(fp));
    put_field(buf, "end", nullptr, 0);
    const Blob blob{0, 0, 0};
    put(buf, &blob, sizeof(blob));
    return buf;
}

// A field of a snapshot buffer: name and data spans.
struct Span {
    BinField f;
    const char* name;
    const char* data;
};

// Reads the field at *pos of buf (after the header); false at the end of buf.
// Sizes are checked against the bytes that remain (a crafted size cannot wrap
// the position), and the name must end with its NUL inside its span.
bool next_field(const std::string& buf, size_t* pos, Span* s) {
    if (*pos > buf.size() || buf.size() - *pos < sizeof(BinField)) return false;
    memcpy(&s->f, buf.data() + *pos, sizeof(BinField));
    const size_t at = *pos + sizeof(BinField);
    const size_t left = buf.size() - at;
    if (s->f.size_name == 0 || s->f.size_name > left) return false;
    if (memchr(buf.data() + at, 0, s->f.size_name) == nullptr) return false;
    if (s->f.size_data > left - s->f.size_name) return false;
    s->name = buf.data() + at;
    s->data = s->name + s->f.size_name;
    *pos = at + s->f.size_name + s->f.size_data;
    return true;
}

bool particles_differ(const char* a, const char* b, size_t bytes) {
    // particle.c:79-94: every double and the name pointer; not sim or ap
    for (size_t off = 0; off + sizeof(struct reb_particle) <= bytes; off += sizeof(struct reb_particle)) {
        struct reb_particle p, q;
        memcpy(&p, a + off, sizeof(p));
        memcpy(&q, b + off, sizeof(q));
        if (p.x != q.x || p.y != q.y || p.z != q.z || p.vx != q.vx || p.vy != q.vy || p.vz != q.vz ||
            p.ax != q.ax || p.ay != q.ay || p.az != q.az || p.m != q.m || p.r != q.r || p.name != q.name)
            return true;
    }
    return false;
}

// binarydata.c:259-488 (stream output): the fields of buf2 that differ from
// or are missing in buf1, and the names of buf1's fields missing in buf2.
std::string diff(const std::string& buf1, const std::string& buf2) {
    std::string out;
    auto find = [](const std::string& buf, const char* name, Span* s) {
        size_t pos = 64;
        Span t;
        while (next_field(buf, &pos, &t)) {
            if (strcmp(t.name, "end") == 0) return false;
            if (strcmp(t.name, name) == 0) {
                *s = t;
                return true;
            }
        }
        return false;
    };
    size_t pos1 = 64;
    Span a;
    while (next_field(buf1, &pos1, &a) && strcmp(a.name, "end") != 0) {
        Span b;
        if (!find(buf2, a.name, &b)) {
            put(out, &a.f, sizeof(BinField));
            put(out, a.name, a.f.size_name);
            continue;
        }
        bool differ = a.f.size_data != b.f.size_data;
        if (!differ) {
            if (strcmp(a.name, "particles") == 0)
                differ = particles_differ(a.data, b.data, a.f.size_data);
            else
                differ = memcmp(a.data, b.data, a.f.size_data) != 0;
        }
        if (differ) {
            put(out, &b.f, sizeof(BinField));
            put(out, b.name, b.f.size_name);
            put(out, b.data, b.f.size_data);
        }
    }
    size_t pos2 = 64;
    Span b;
    while (next_field(buf2, &pos2, &b) && strcmp(b.name, "end") != 0) {
        Span t;
        if (find(buf1, b.name, &t)) continue;
        put(out, &b.f, sizeof(BinField));
        put(out, b.name, b.f.size_name);
        put(out, b.data, b.f.size_data);
    }
    return out;
}

bool read_file(const char* filename, std::string* out) {
    FILE* f = fopen(filename, "rb");
    if (!f) return false;
    std::string buf;
    char chunk[1 << 16];
    size_t n;
    while ((n = fread(chunk, 1, sizeof(chunk), f)) > 0) buf.append(chunk, n);
    fclose(f);
    *out = std::move(buf);
    return true;
}

// Offsets of the snapshots of an archive (simulationarchive.c:115-318): the
// first snapshot at 0, each later one just past the previous blob.
std::vector<size_t> snapshot_offsets(const std::string& buf, bool* corrupt) {
    std::vector<size_t> off;
    *corrupt = false;
    if (buf.size() < 64) {
        *corrupt = true;
        return off;
    }
    size_t pos = 0;
    for (;;) {
        const size_t start = pos;
        if (start == 0) pos = 64;
        Span s;
        bool ended = false;
        while (next_field(buf, &pos, &s)) {
            if (strcmp(s.name, "end") == 0) {
                ended = true;
                break;
            }
        }
        if (!ended || pos + sizeof(Blob) > buf.size()) {
            *corrupt = true;
            break;
        }
        Blob blob;
        memcpy(&blob, buf.data() + pos, sizeof(blob));
        pos += sizeof(blob);
        if (!off.empty() && (int64_t)blob.offset_prev + (int64_t)sizeof(Blob) != (int64_t)(pos - start)) {
            *corrupt = true;
            break;
        }
        off.push_back(start);
        if (blob.offset_next == 0) break;
    }
    return off;
}

enum Warn : unsigned {
    W_VERSION = 1,
    W_POINTERS = 2,
    W_UNKNOWN = 4,
    W_CORRUPT = 8,
    W_INTEGRATOR = 16,
};

// binarydata.c:680-879: applies the fields from `pos` up to the "end" field.
void input_fields(struct reb_simulation* r, const std::string& buf, size_t pos, unsigned* warnings) {
    for (;;) {
        if (pos + sizeof(BinField) > buf.size()) return;
        BinField f;
        memcpy(&f, buf.data() + pos, sizeof(f));
        if (f.size_name == kHeaderTag) {
            char cur[64];
            snprintf(cur, sizeof(cur), "REBOUND Binary File. Version: %s", reb_version_str);
            if (strncmp(buf.data() + pos, cur, strlen(cur) + 1) != 0) *warnings |= W_VERSION;
            pos += 64;
            continue;
        }
        Span s;
        if (!next_field(buf, &pos, &s) || s.f.size_name > REB_STRING_SIZE_MAX) {
            *warnings |= W_CORRUPT;
            return;
        }
        const char* name = s.name;
        if (strcmp(name, "end") == 0) return;
        if (strcmp(name, "functionpointers") == 0) {
            int fp = 0;
            memcpy(&fp, s.data, std::min<size_t>(sizeof(fp), s.f.size_data));
            if (fp) *warnings |= W_POINTERS;
            continue;
        }
        if (strcmp(name, "integrator.name") == 0) {
            std::string iname(s.data, s.f.size_data ? s.f.size_data - 1 : 0);
            reb_simulation_set_integrator(r, iname.c_str());
            continue;
        }
        const Desc* d = nullptr;
        char* base = nullptr;
        if (strncmp(name, "integrator.", 11) == 0) {
            const char* sub = strchr(name + 11, '.');
            if (!sub || !r->integrator.name || strncmp(name + 11, r->integrator.name, sub - name - 11) != 0 ||
                strlen(r->integrator.name) != (size_t)(sub - name - 11)) {
                *warnings |= W_CORRUPT;
                return;
            }
            static thread_local std::vector<Desc> ilist;
            ilist = integrator_fields(r->integrator.name);
            for (const Desc& e : ilist)
                if (strcmp(e.name, sub + 1) == 0) d = &e;
            base = static_cast<char*>(r->integrator.state);
            if (!d || !base) {
                *warnings |= W_UNKNOWN;
                continue;
            }
        } else {
            for (const Desc& e : kSimFields)
                if (strcmp(e.name, name) == 0) d = &e;
            base = reinterpret_cast<char*>(r);
            if (!d) {
                *warnings |= W_UNKNOWN;
                continue;
            }
        }
        char* at = base + d->offset;
        switch (d->type) {
            case T::POINTER: {
                void** p = reinterpret_cast<void**>(at);
                void* q = realloc(*p, s.f.size_data ? s.f.size_data : 1);
                if (!q) {
                    reb_simulation_error(r, "Out of memory while reading a binary file.");
                    return;
                }
                *p = q;
                memcpy(q, s.data, s.f.size_data);
                if (d->offset_n != kNoCount) *reinterpret_cast<size_t*>(base + d->offset_n) = s.f.size_data / d->element;
                break;
            }
            case T::STRING: {
                char** p = reinterpret_cast<char**>(at);
                char* q = static_cast<char*>(realloc(*p, s.f.size_data + 1));
                memcpy(q, s.data, s.f.size_data);
                q[s.f.size_data] = 0;
                *p = q;
                break;
            }
            case T::CHARP_LIST: {
                size_t* n = reinterpret_cast<size_t*>(base + d->offset_n);
                char*** list = reinterpret_cast<char***>(at);
                size_t q = 0;
                while (q < s.f.size_data) {
                    // each entry: a NUL-terminated name, then the pointer it had
                    const char* str = s.data + q;
                    const size_t room = s.f.size_data - q;
                    const size_t slen = strnlen(str, room);
                    if (slen == room || room - slen - 1 < sizeof(char*)) {
                        *warnings |= W_CORRUPT;
                        return;
                    }
                    const size_t len = slen + 1 + sizeof(char*);
                    (*n)++;
                    *list = static_cast<char**>(realloc(*list, sizeof(char*) * (*n)));
                    (*list)[*n - 1] = static_cast<char*>(malloc(len));
                    memcpy((*list)[*n - 1], str, len);
                    q += len;
                }
                break;
            }
            default:
                memcpy(at, s.data, std::min<size_t>(simple_size(d->type), s.f.size_data));
                break;
        }
    }
}

// binarydata.c:851-878: the state the fields leave behind.
void finish_input(struct reb_simulation* r, unsigned* warnings) {
    if (r->N && !r->particles) {
        // a damaged file gave the count but not the bodies
        *warnings |= W_CORRUPT;
        r->N = 0;
    }
    if (r->N_var && !r->particles_var) {
        *warnings |= W_CORRUPT;
        r->N_var = 0;
    }
    if (r->N_var_config && !r->var_config) {
        *warnings |= W_CORRUPT;
        r->N_var_config = 0;
    }
    for (size_t l = 0; l < r->N_var_config; l++) r->var_config[l].sim = r;
    for (size_t l = 0; l < r->N_var; l++) {
        r->particles_var[l].ap = nullptr;
        r->particles_var[l].sim = r;
    }
    r->N_allocated = r->N;
    for (size_t l = 0; l < r->N; l++) {
        r->particles[l].ap = nullptr;
        r->particles[l].sim = r;
        if (r->particles[l].name) {
            bool found = false;
            for (size_t k = 0; k < r->N_name_list; k++) {
                const char* orig = *reinterpret_cast<char* const*>(r->name_list[k] + strlen(r->name_list[k]) + 1);
                if (r->particles[l].name == orig) {
                    r->particles[l].name = r->name_list[k];
                    found = true;
                }
            }
            if (!found) {
                reb_simulation_warning(r, "A name for a particle was not stored in the Simulationarchive.");
                r->particles[l].name = nullptr;
            }
        }
    }
    mark_host_newer(r);
}

void report(struct reb_simulation* r, unsigned w) {
    if (w & W_VERSION)
        reb_simulation_warning(r, "Binary file was saved with a different version of REBOUND. Binary format might have changed.");
    if (w & W_POINTERS)
        reb_simulation_warning(r, "You have to reset function pointers after creating a reb_simulation struct with a binary file.");
    if (w & W_UNKNOWN) reb_simulation_warning(r, "Unknown field found in binary file.");
    if (w & W_CORRUPT)
        reb_simulation_warning(r, "The binary file seems to be corrupted. An attempt has been made to read the uncorrupted parts of it.");
}

// The reference's REB_BINARYDATA_* codes (binarydata.h:31-45), which the
// *_with_messages entry points report (its Python package maps them to
// exceptions and warnings).
enum RefCode : int {
    RC_NOFILE = 1,
    RC_VERSION = 2,
    RC_POINTERS = 4,
    RC_FILENOTOPEN = 16,
    RC_OUTOFRANGE = 32,
    RC_SEEK = 64,
    RC_FIELD_UNKNOWN = 128,
    RC_CORRUPTFILE = 512,
    RC_OLD = 1024,
};

int ref_codes(unsigned w) {
    int c = 0;
    if (w & W_VERSION) c |= RC_VERSION;
    if (w & W_POINTERS) c |= RC_POINTERS;
    if (w & W_UNKNOWN) c |= RC_FIELD_UNKNOWN;
    if (w & W_CORRUPT) c |= RC_CORRUPTFILE;
    return c;
}

// binarydata.c:881-930: messages for the codes; NULL when an error ends the
// read (the simulation is freed).
struct reb_simulation* process_codes(struct reb_simulation* r, int c) {
    auto fail = [&](const char* msg) -> struct reb_simulation* {
        reb_simulation_error(r, msg);
        if (r) reb_simulation_free(r);
        return nullptr;
    };
    if (c & RC_NOFILE) return fail("Cannot read binary file. Check filename and file contents.");
    if (c & RC_VERSION)
        reb_simulation_warning(r, "Binary file was saved with a different version of REBOUND. Binary format might have changed.");
    if (c & RC_POINTERS)
        reb_simulation_warning(r, "You have to reset function pointers after creating a reb_simulation struct with a binary file.");
    if (c & RC_FILENOTOPEN) return fail("Error while reading binary file (file was not open).");
    if (c & RC_OUTOFRANGE) return fail("Index out of range.");
    if (c & RC_SEEK) return fail("Error while trying to seek file.");
    if (c & RC_FIELD_UNKNOWN) reb_simulation_warning(r, "Unknown field found in binary file.");
    if (c & RC_OLD)
        return fail("Reading old Simulationarchives (version < 2) is no longer supported. If you need to read such an "
                    "archive, use a REBOUND version <= 3.26.3");
    if (c & RC_CORRUPTFILE)
        reb_simulation_warning(r, "The binary file seems to be corrupted. An attempt has been made to read the uncorrupted parts of it.");
    return r;
}

// Readable value of a field for the diff printout: doubles and integers by
// value, other data by size.
std::string field_text(const Span& s) {
    char tmp[64];
    if (strcmp(s.name, "particles") == 0) {
        std::string out;
        for (size_t off = 0; off + sizeof(struct reb_particle) <= s.f.size_data; off += sizeof(struct reb_particle)) {
            struct reb_particle p;
            memcpy(&p, s.data + off, sizeof(p));
            snprintf(tmp, sizeof(tmp), "%s(%.17g %.17g %.17g", off ? " " : "", p.x, p.y, p.z);
            out += tmp;
            snprintf(tmp, sizeof(tmp), " %.17g %.17g %.17g m %.17g)", p.vx, p.vy, p.vz, p.m);
            out += tmp;
        }
        return out;
    }
    const Desc* d = nullptr;
    for (const Desc& e : kSimFields)
        if (strcmp(e.name, s.name) == 0) d = &e;
    if (d && d->type == T::DOUBLE && s.f.size_data == sizeof(double)) {
        double v;
        memcpy(&v, s.data, sizeof(v));
        snprintf(tmp, sizeof(tmp), "%.17g", v);
        return tmp;
    }
    if (d && (d->type == T::INT || d->type == T::UINT || d->type == T::UINT32) && s.f.size_data == 4) {
        int32_t v;
        memcpy(&v, s.data, sizeof(v));
        snprintf(tmp, sizeof(tmp), "%d", v);
        return tmp;
    }
    if (d && (d->type == T::INT64 || d->type == T::UINT64 || d->type == T::SIZE_T) && s.f.size_data == 8) {
        int64_t v;
        memcpy(&v, s.data, sizeof(v));
        snprintf(tmp, sizeof(tmp), "%lld", (long long)v);
        return tmp;
    }
    snprintf(tmp, sizeof(tmp), "(%llu bytes)", (unsigned long long)s.f.size_data);
    return tmp;
}

// binarydata.c:259-488 (print and no-output modes): whether the two
// snapshots differ, ignoring fields whose names start with "walltime", and
// the differences as text.
bool compare(const std::string& buf1, const std::string& buf2, std::string* text) {
    bool different = false;
    auto find = [](const std::string& buf, const char* name, Span* sp) {
        size_t pos = 64;
        Span t;
        while (next_field(buf, &pos, &t)) {
            if (strcmp(t.name, "end") == 0) return false;
            if (strcmp(t.name, name) == 0) {
                *sp = t;
                return true;
            }
        }
        return false;
    };
    if (buf1.size() >= 64 && buf2.size() >= 64 && memcmp(buf1.data(), buf2.data(), 64) != 0)
        *text += "Header in binary files are different.\n";
    size_t pos1 = 64;
    Span a;
    while (next_field(buf1, &pos1, &a) && strcmp(a.name, "end") != 0) {
        Span b;
        if (!find(buf2, a.name, &b)) {
            different = true;
            *text += std::string(a.name) + ":\n< " + field_text(a) + "\n";
            continue;
        }
        bool differ = a.f.size_data != b.f.size_data;
        if (!differ) {
            if (strcmp(a.name, "particles") == 0)
                differ = particles_differ(a.data, b.data, a.f.size_data);
            else
                differ = memcmp(a.data, b.data, a.f.size_data) != 0;
        }
        if (differ) {
            if (strncmp(a.name, "walltime", 8) != 0) different = true;
            *text += std::string(a.name) + ":\n< " + field_text(a) + "\n---\n> " + field_text(b) + "\n";
        }
    }
    size_t pos2 = 64;
    Span b;
    while (next_field(buf2, &pos2, &b) && strcmp(b.name, "end") != 0) {
        Span t;
        if (find(buf1, b.name, &t)) continue;
        different = true;
        *text += std::string(b.name) + ":\n> " + field_text(b) + "\n";
    }
    return different;
}

}  // namespace

// The host state a snapshot must see: particles (and WHFast's p_jh) current.
bool prepare_host(struct reb_simulation* r) {
    if (!ensure_host(r)) return false;
    if (r->integrator.name && strcmp(r->integrator.name, "whfast") == 0) whfast_refresh_host(r);
    return true;
}

// simulationarchive.c:377-409
void archive_heartbeat(struct reb_simulation* r) {
    if (!r->simulationarchive_filename) return;
    int modes = 0;
    if (r->simulationarchive_auto_interval != 0) modes++;
    if (r->simulationarchive_auto_walltime != 0.) modes++;
    if (r->simulationarchive_auto_step != 0) modes++;
    if (modes > 1)
        reb_simulation_error(r, "Only use one of simulationarchive_auto_interval, simulationarchive_auto_walltime, or "
                                "simulationarchive_auto_step");
    if (r->simulationarchive_auto_interval != 0.) {
        const double sign = r->dt > 0. ? 1. : -1;
        if (sign * r->simulationarchive_next <= sign * r->t) {
            r->simulationarchive_next += sign * r->simulationarchive_auto_interval;
            reb_simulation_save_to_file(r, nullptr);
        }
    }
    if (r->simulationarchive_auto_step != 0.) {
        if (r->simulationarchive_next_step <= r->steps_done) {
            r->simulationarchive_next_step += r->simulationarchive_auto_step;
            reb_simulation_save_to_file(r, nullptr);
        }
    }
    if (r->simulationarchive_auto_walltime != 0.) {
        if (r->simulationarchive_next <= r->walltime) {
            r->simulationarchive_next += r->simulationarchive_auto_walltime;
            reb_simulation_save_to_file(r, nullptr);
        }
    }
}

}  // namespace rbamd

using namespace rbamd;

extern "C" {

// simulationarchive.c:411-615: a new file holds the whole simulation; an
// existing archive gets the diff against its first snapshot appended.
REB_API void reb_simulation_save_to_file(struct reb_simulation* const r, const char* filename) {
    if (r->simulationarchive_version < 5) {
        reb_simulation_error(r, "Writing Simulationarchives with a version < 5 is no longer supported.\n");
        return;
    }
    if (!filename) filename = r->simulationarchive_filename;
    if (!filename) {
        reb_simulation_error(r, "Filename missing.");
        return;
    }
    if (!prepare_host(r)) return;
    struct stat st;
    if (stat(filename, &st) < 0) {
        const std::string buf = serialize(r);
        FILE* of = fopen(filename, "wb");
        if (!of) {
            reb_simulation_error(r, "Can not open file.");
            return;
        }
        fwrite(buf.data(), buf.size(), 1, of);
        fclose(of);
        return;
    }
    std::string file;
    bool corrupt = false;
    if (!read_file(filename, &file)) {
        reb_simulation_error(r, "Can not open file.");
        return;
    }
    const std::vector<size_t> off = snapshot_offsets(file, &corrupt);
    if (off.empty()) {
        reb_simulation_warning(r, "Simulationarchive appears to be corrupted. A recovery attempt has failed. No snapshot has been saved.\n");
        return;
    }
    if (corrupt)
        reb_simulation_warning(r, "Simulationarchive appears to be corrupted. REBOUND will attempt to fix it before appending more snapshots.\n");
    // the first snapshot up to and including its "end" field
    size_t pos = 64, end1 = 0;
    Span s;
    while (next_field(file, &pos, &s))
        if (strcmp(s.name, "end") == 0) {
            end1 = pos;
            break;
        }
    const std::string first = file.substr(0, end1);
    const std::string d = diff(first, serialize(r));
    // the last valid blob: after the last snapshot's "end" field
    size_t last_blob = end1;
    if (off.size() > 1) {
        size_t p = off.back();
        while (next_field(file, &p, &s))
            if (strcmp(s.name, "end") == 0) break;
        last_blob = p;
    }
    Blob blob;
    memcpy(&blob, file.data() + last_blob, sizeof(blob));
    blob.offset_next = (int32_t)(d.size() + sizeof(BinField) + strlen("end") + 1);
    std::string tail;
    put(tail, &blob, sizeof(blob));
    put(tail, d.data(), d.size());
    put_field(tail, "end", nullptr, 0);
    Blob next{blob.index + 1, blob.offset_next, 0};
    put(tail, &next, sizeof(next));
    FILE* of = fopen(filename, "r+b");
    if (!of) {
        reb_simulation_error(r, "Can not open file.");
        return;
    }
    fseek(of, (long)last_blob, SEEK_SET);
    fwrite(tail.data(), tail.size(), 1, of);
    fclose(of);
    if (last_blob + tail.size() < file.size()) {
        // a cut-off snapshot followed the last valid one: drop it
        if (truncate(filename, (off_t)(last_blob + tail.size())) != 0)
            reb_simulation_warning(r, "Could not truncate the Simulationarchive after its last valid snapshot.");
    }
}

// simulationarchive.c:95-113 and :45-84
REB_API struct reb_simulation* reb_simulation_create_from_file(char* filename, int64_t snapshot) {
    std::string file;
    if (!filename || !read_file(filename, &file)) return nullptr;
    struct reb_simulation* r = reb_simulation_create();
    bool corrupt = false;
    const std::vector<size_t> off = snapshot_offsets(file, &corrupt);
    unsigned w = corrupt ? (unsigned)W_CORRUPT : 0u;
    if (off.empty()) {
        reb_simulation_error(r, "Cannot read binary file. Check filename and file contents.");
        reb_simulation_free(r);
        return nullptr;
    }
    const int64_t nblobs = (int64_t)off.size();
    if (snapshot < 0) snapshot += nblobs;
    if (snapshot < 0 || snapshot >= nblobs) {
        reb_simulation_error(r, "Index out of range.");
        reb_simulation_free(r);
        return nullptr;
    }
    free(r->simulationarchive_filename);
    r->simulationarchive_filename = nullptr;
    input_fields(r, file, 0, &w);
    if (snapshot > 0) input_fields(r, file, off[(size_t)snapshot], &w);
    finish_input(r, &w);
    report(r, w);
    return r;
}

// simulationarchive.c:617-666
static int set_archive_filename(struct reb_simulation* r, const char* filename) {
    if (!filename) {
        reb_simulation_error(r, "Filename missing.");
        return -1;
    }
    struct stat st;
    if (stat(filename, &st) == 0)
        reb_simulation_warning(r, "File in use for Simulationarchive already exists. Snapshots will be appended.");
    free(r->simulationarchive_filename);
    r->simulationarchive_filename = strdup(filename);
    return 0;
}

REB_API void reb_simulation_save_to_file_interval(struct reb_simulation* const r, const char* filename,
                                                  double interval) {
    if (set_archive_filename(r, filename) < 0) return;
    if (r->simulationarchive_auto_interval != interval) {
        r->simulationarchive_auto_interval = interval;
        r->simulationarchive_next = r->t;
    }
}

REB_API void reb_simulation_save_to_file_walltime(struct reb_simulation* const r, const char* filename,
                                                  double walltime) {
    if (set_archive_filename(r, filename) < 0) return;
    r->simulationarchive_auto_walltime = walltime;
    r->simulationarchive_next = r->walltime;
}

REB_API void reb_simulation_save_to_file_step(struct reb_simulation* const r, const char* filename, uint64_t step) {
    if (set_archive_filename(r, filename) < 0) return;
    if (r->simulationarchive_auto_step != step) {
        r->simulationarchive_auto_step = step;
        r->simulationarchive_next_step = r->steps_done;
    }
}

// ------------------------------------------------------------ comparison and streams
// simulation.c:691-709
REB_API int reb_simulation_diff(struct reb_simulation* r1, struct reb_simulation* r2, int output_option) {
    if (output_option != 0 && output_option != 1) {
        reb_simulation_error(r1, "Invalid output_option in reb_simulation_diff().");
        return -1;
    }
    if (!prepare_host(r1) || !prepare_host(r2)) return -1;
    std::string text;
    const bool different = compare(serialize(r1), serialize(r2), &text);
    if (output_option == 1) {
        fputs(text.c_str(), stdout);
        fflush(stdout);
    }
    return different ? 1 : 0;
}

// simulation.c:618-634: the differences as a string the caller frees with reb_free.
REB_API char* reb_simulation_diff_char(struct reb_simulation* r1, struct reb_simulation* r2) {
    std::string text;
    if (prepare_host(r1) && prepare_host(r2)) compare(serialize(r1), serialize(r2), &text);
    return strdup(text.c_str());
}

// binarydata.c:582-606: the simulation as one snapshot in memory (malloc'd).
REB_API void reb_binarydata_simulation_to_stream(struct reb_simulation* r, char** bufp, size_t* sizep) {
    *bufp = nullptr;
    *sizep = 0;
    if (!prepare_host(r)) return;
    const std::string buf = serialize(r);
    *bufp = static_cast<char*>(malloc(buf.size()));
    if (!*bufp) return;
    memcpy(*bufp, buf.data(), buf.size());
    *sizep = buf.size();
}

// simulation.c:711-717 through the binary format (reb_simulation_copy keeps
// its direct struct copy): r_copy is a freshly created simulation.
REB_API void reb_simulation_copy_with_messages(struct reb_simulation* r_copy, struct reb_simulation* r,
                                               int* warnings) {
    unsigned w = 0;
    if (prepare_host(r)) {
        input_fields(r_copy, serialize(r), 0, &w);
        finish_input(r_copy, &w);
        dev_of(r_copy)->precision = dev_of(r)->precision;
        dev_of(r_copy)->cor_constant = dev_of(r)->cor_constant;
    }
    if (warnings) *warnings |= ref_codes(w);
}

// ------------------------------------------------------------ archive objects
// simulationarchive.c:115-375.  The index is built once from the whole file
// (the snapshot offsets and each snapshot's time); the stream stays open and
// each snapshot read takes only its own bytes and the first snapshot's.

static void archive_index(struct reb_simulationarchive* sa, const std::string& file, int* codes) {
    bool corrupt = false;
    const std::vector<size_t> off = snapshot_offsets(file, &corrupt);
    // header "REBOUND Binary File. Version: major.minor.patch"
    const char* colon = static_cast<const char*>(memchr(file.data(), ':', std::min<size_t>(file.size(), 64)));
    if (colon && sscanf(colon + 1, "%d.%d.%d", &sa->reb_version_major, &sa->reb_version_minor,
                        &sa->reb_version_patch) != 3)
        *codes |= RC_CORRUPTFILE;
    if (file.size() >= 64) {
        char cur[64];
        snprintf(cur, sizeof(cur), "REBOUND Binary File. Version: %s", reb_version_str);
        if (strncmp(file.data(), cur, strlen(cur) + 1) != 0) *codes |= RC_VERSION;
    }
    size_t pos = 64;
    Span f;
    while (next_field(file, &pos, &f) && strcmp(f.name, "end") != 0) {
        auto get = [&](void* dst, size_t n) { memcpy(dst, f.data, std::min<size_t>(n, f.f.size_data)); };
        if (strcmp(f.name, "simulationarchive_version") == 0) get(&sa->version, sizeof(sa->version));
        if (strcmp(f.name, "simulationarchive_auto_interval") == 0) get(&sa->auto_interval, sizeof(double));
        if (strcmp(f.name, "simulationarchive_auto_walltime") == 0) get(&sa->auto_walltime, sizeof(double));
        if (strcmp(f.name, "simulationarchive_auto_step") == 0) get(&sa->auto_step, sizeof(uint64_t));
    }
    if (sa->version < 5) {
        *codes |= RC_OLD;
        return;
    }
    if (off.empty()) {
        *codes |= RC_SEEK;
        return;
    }
    if (corrupt) *codes |= RC_CORRUPTFILE;
    sa->nblobs = (int64_t)off.size();
    sa->offset = static_cast<uint64_t*>(calloc(off.size(), sizeof(uint64_t)));
    sa->t = static_cast<double*>(calloc(off.size(), sizeof(double)));
    double t = 0.;
    for (size_t k = 0; k < off.size(); k++) {
        sa->offset[k] = off[k];
        size_t p = k == 0 ? 64 : off[k];
        while (next_field(file, &p, &f) && strcmp(f.name, "end") != 0)
            if (strcmp(f.name, "t") == 0 && f.f.size_data == sizeof(double)) memcpy(&t, f.data, sizeof(double));
        sa->t[k] = t;
    }
}

static bool stream_bytes(FILE* inf, std::string* out) {
    out->clear();
    if (fseek(inf, 0, SEEK_SET) != 0) return false;
    char chunk[1 << 16];
    size_t n;
    while ((n = fread(chunk, 1, sizeof(chunk), inf)) > 0) out->append(chunk, n);
    return true;
}

REB_API struct reb_simulationarchive* reb_simulationarchive_create_from_file_with_messages(const char* filename,
                                                                                         int* warnings) {
    auto* sa = static_cast<struct reb_simulationarchive*>(calloc(1, sizeof(struct reb_simulationarchive)));
    int codes = 0;
    sa->inf = filename ? fopen(filename, "rb") : nullptr;
    sa->filename = filename ? strdup(filename) : nullptr;
    std::string file;
    if (!sa->inf || !stream_bytes(sa->inf, &file))
        codes |= RC_NOFILE;
    else
        archive_index(sa, file, &codes);
    if ((codes & (RC_OLD | RC_SEEK)) && sa->inf) {
        fclose(sa->inf);
        sa->inf = nullptr;
    }
    if (warnings) *warnings |= codes;
    return sa;
}

REB_API struct reb_simulationarchive* reb_simulationarchive_create_from_buffer_with_messages(char* buffer,
                                                                                           size_t size,
                                                                                           int* warnings) {
    auto* sa = static_cast<struct reb_simulationarchive*>(calloc(1, sizeof(struct reb_simulationarchive)));
    int codes = 0;
    sa->inf = (buffer && size) ? fmemopen(buffer, size, "rb") : nullptr;
    if (!sa->inf)
        codes |= RC_NOFILE;
    else
        archive_index(sa, std::string(buffer, size), &codes);
    if ((codes & (RC_OLD | RC_SEEK)) && sa->inf) {
        fclose(sa->inf);
        sa->inf = nullptr;
    }
    if (warnings) *warnings |= codes;
    return sa;
}

REB_API struct reb_simulationarchive* reb_simulationarchive_create_from_file(const char* filename) {
    int codes = 0;
    struct reb_simulationarchive* sa = reb_simulationarchive_create_from_file_with_messages(filename, &codes);
    if (codes & RC_NOFILE) {
        // a missing file gives NULL without a message (simulationarchive.c:345-356)
        reb_simulationarchive_free(sa);
        return nullptr;
    }
    if (codes & (RC_OLD | RC_SEEK)) {
        process_codes(nullptr, codes);
        reb_simulationarchive_free(sa);
        return nullptr;
    }
    process_codes(nullptr, codes);
    return sa;
}

REB_API void reb_simulationarchive_free(struct reb_simulationarchive* sa) {
    if (!sa) return;
    if (sa->inf) fclose(sa->inf);
    free(sa->filename);
    free(sa->t);
    free(sa->offset);
    free(sa);
}

// simulationarchive.c:45-84: r is a freshly created simulation.
REB_API void reb_simulation_init_from_simulationarchive_with_messages(struct reb_simulation* r,
                                                                     struct reb_simulationarchive* sa,
                                                                     int64_t snapshot, int* warnings) {
    int codes = 0;
    if (!sa || !sa->inf) {
        codes |= RC_FILENOTOPEN;
    } else {
        if (snapshot < 0) snapshot += sa->nblobs;
        if (snapshot < 0 || snapshot >= sa->nblobs) codes |= RC_OUTOFRANGE;
    }
    if (codes) {
        if (warnings) *warnings |= codes;
        return;
    }
    std::string file;
    if (!stream_bytes(sa->inf, &file)) {
        if (warnings) *warnings |= RC_SEEK;
        return;
    }
    free(r->simulationarchive_filename);
    r->simulationarchive_filename = nullptr;
    const size_t end0 = sa->nblobs > 1 ? (size_t)sa->offset[1] : file.size();
    unsigned w = 0;
    input_fields(r, file.substr(0, end0), 0, &w);
    if (snapshot > 0) {
        const size_t a = (size_t)sa->offset[snapshot];
        const size_t b = snapshot + 1 < sa->nblobs ? (size_t)sa->offset[snapshot + 1] : file.size();
        if (a > file.size() || b > file.size() || b < a) {
            if (warnings) *warnings |= RC_SEEK;
            return;
        }
        input_fields(r, file.substr(a, b - a), 0, &w);
    }
    finish_input(r, &w);
    if (warnings) *warnings |= ref_codes(w);
}

// simulationarchive.c:86-93
REB_API struct reb_simulation* reb_simulation_create_from_simulationarchive(struct reb_simulationarchive* sa,
                                                                           int64_t snapshot) {
    if (!sa) return nullptr;
    struct reb_simulation* r = reb_simulation_create();
    int codes = 0;
    reb_simulation_init_from_simulationarchive_with_messages(r, sa, snapshot, &codes);
    return process_codes(r, codes);
}

}  // extern "C"
