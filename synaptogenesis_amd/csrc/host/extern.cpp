// The reference-compatible C-ABI (src/core/extern.cpp:12-375).
#include "syngen.h"

#include <hip/hip_runtime.h>

#include <cstdlib>
#include <cstring>
#include <fstream>
#include <mutex>
#include <set>
#include <sstream>

#include "callbacks.h"
#include "engine.h"
#include "network.h"
#include "props.h"
#include "state.h"

using namespace sg;

namespace {

// Handles carry a tag so destroy() can tell what it is given.
enum Tag : uint32_t { TAG_PROPS = 0x50524f50, TAG_NET = 0x4e455457,
                      TAG_ENV = 0x454e5649, TAG_STATE = 0x53544154 };

struct NetHandle {
    uint32_t tag = TAG_NET;
    std::unique_ptr<Network> net;
};
struct EnvHandle {
    uint32_t tag = TAG_ENV;
    Props props;
};
struct StateHandle {
    uint32_t tag = TAG_STATE;
    std::unique_ptr<State> state;
};
// Props handles are bare Props objects; destroy() tells handles apart with
// a registry of the live network / environment / state handles.
std::mutex g_handles_mu;
std::set<void *> g_handles;

template <class T>
T *track(T *h) {
    std::lock_guard<std::mutex> g(g_handles_mu);
    g_handles.insert(h);
    return h;
}

template <class T>
void untrack_delete(T *h) {
    if (!h) return;
    {
        std::lock_guard<std::mutex> g(g_handles_mu);
        g_handles.erase(h);
    }
    delete h;
}

bool is_handle(void *p) {
    std::lock_guard<std::mutex> g(g_handles_mu);
    return g_handles.count(p) > 0;
}

bool g_suppress = false, g_warnings = true, g_debug = false;

ARRAY null_array() { return ARRAY{0, VOID_POINTER, nullptr, false}; }

ARRAY string_array(const std::vector<std::string> &strings) {
    if (strings.empty()) return null_array();
    char **pp = (char **)malloc(strings.size() * sizeof(char *));
    for (size_t i = 0; i < strings.size(); ++i) {
        pp[i] = (char *)malloc(strings[i].size() + 1);
        std::memcpy(pp[i], strings[i].c_str(), strings[i].size() + 1);
    }
    return ARRAY{(int)strings.size(), STRING_POINTER, pp, true};
}

ARRAY view_array(Array *a) {
    // Output words are reported as FLOAT_POINTER (extern.cpp:12-22).
    POINTER_TYPE t = a->elem == Elem::I32 ? INT_POINTER : FLOAT_POINTER;
    return ARRAY{(int)a->size(), t, a->host.data(), false};
}

std::string read_file(const char *path) {
    std::ifstream f(path);
    if (!f) throw Error(std::string("Could not open file: ") + path);
    std::stringstream ss;
    ss << f.rdbuf();
    return ss.str();
}

Props *P(PROPS p) { return static_cast<Props *>(p); }

}  // namespace

extern "C" {

void free_array(ARRAY arr) {
    if (arr.owner) free(arr.data);
}

void free_array_deep(ARRAY arr) {
    if (arr.owner) {
        for (int i = 0; i < arr.size; ++i) free(((void **)arr.data)[i]);
        free(arr.data);
    }
}

PROPS create_properties(void) { return new Props(); }

void add_property(PROPS properties, char *key, char *val) {
    P(properties)->set(key, val);
}

void add_child(PROPS properties, char *key, PROPS child) {
    P(properties)->set_child(key, *P(child));
}

void add_to_array(PROPS properties, char *key, char *val) {
    P(properties)->add_to_array(key, val);
}

void add_to_child_array(PROPS properties, char *key, PROPS props) {
    P(properties)->add_to_child_array(key, *P(props));
}

ARRAY get_keys(PROPS p) { return string_array(P(p)->keys()); }
ARRAY get_child_keys(PROPS p) { return string_array(P(p)->child_keys()); }
ARRAY get_array_keys(PROPS p) { return string_array(P(p)->array_keys()); }
ARRAY get_child_array_keys(PROPS p) {
    return string_array(P(p)->child_array_keys());
}

const char *get_property(PROPS p, char *key) {
    return P(p)->has(key) ? P(p)->get_c_str(key) : nullptr;
}

PROPS get_child(PROPS p, char *key) { return P(p)->get_child(key); }

ARRAY get_array(PROPS p, char *key) {
    if (!P(p)->has_array(key)) return null_array();
    return string_array(P(p)->get_array(key));
}

ARRAY get_child_array(PROPS p, char *key) {
    auto arr = P(p)->get_child_array(key);
    if (arr.empty()) return null_array();
    void **pp = (void **)malloc(arr.size() * sizeof(void *));
    for (size_t i = 0; i < arr.size(); ++i) pp[i] = arr[i];
    return ARRAY{(int)arr.size(), PROPS_POINTER, pp, true};
}

ENVIRONMENT create_environment(PROPS properties) {
    if (!properties) return nullptr;
    try {
        EnvHandle *e = new EnvHandle();
        e->props = *P(properties);
        return track(e);
    } catch (...) {
        return nullptr;
    }
}

ENVIRONMENT load_env(char *filename) {
    try {
        std::unique_ptr<EnvHandle> e(new EnvHandle());
        e->props = parse_json(read_file(filename));
        return track(e.release());
    } catch (...) {
        return nullptr;
    }
}

bool save_env(ENVIRONMENT env, char *filename) {
    if (!env || !filename) return false;
    try {
        std::ofstream f(filename);
        f << write_json(static_cast<EnvHandle *>(env)->props);
        return (bool)f;
    } catch (...) {
        return false;
    }
}

NETWORK create_network(PROPS properties) {
    if (!properties) return nullptr;
    try {
        std::unique_ptr<NetHandle> n(new NetHandle());
        n->net.reset(new Network(*P(properties)));
        return track(n.release());
    } catch (const std::exception &ex) {
        if (!g_suppress) fprintf(stderr, "syngen: %s\n", ex.what());
        return nullptr;
    } catch (...) {
        return nullptr;
    }
}

NETWORK load_net(char *filename) {
    try {
        std::unique_ptr<NetHandle> n(new NetHandle());
        n->net.reset(new Network(parse_json(read_file(filename))));
        return track(n.release());
    } catch (...) {
        return nullptr;
    }
}

bool save_net(NETWORK network, char *filename) {
    if (!network || !filename) return false;
    try {
        std::ofstream f(filename);
        f << write_json(static_cast<NetHandle *>(network)->net->config());
        return (bool)f;
    } catch (...) {
        return false;
    }
}

STATE build_state(NETWORK net) {
    if (!net) return nullptr;
    try {
        std::unique_ptr<StateHandle> s(new StateHandle());
        s->state.reset(new State(static_cast<NetHandle *>(net)->net.get()));
        return track(s.release());
    } catch (...) {
        return nullptr;
    }
}

STATE build_load_state(NETWORK net, char *filename) {
    if (!net || !filename) return nullptr;
    try {
        std::unique_ptr<StateHandle> s(new StateHandle());
        s->state.reset(new State(static_cast<NetHandle *>(net)->net.get()));
        s->state->load(filename);
        return track(s.release());
    } catch (...) {
        return nullptr;
    }
}

bool load_state(STATE state, char *filename) {
    if (!state || !filename) return false;
    try {
        static_cast<StateHandle *>(state)->state->load(filename);
        return true;
    } catch (...) {
        return false;
    }
}

bool save_state(STATE state, char *filename) {
    if (!state || !filename) return false;
    try {
        static_cast<StateHandle *>(state)->state->save(filename);
        return true;
    } catch (...) {
        return false;
    }
}

ARRAY get_neuron_data(STATE state, char *structure_name, char *layer_name,
                      char *key) {
    if (!state || !structure_name || !layer_name || !key) return null_array();
    try {
        State *st = static_cast<StateHandle *>(state)->state.get();
        Structure *s = st->network()->structure(structure_name);
        if (!s) return null_array();
        Layer *l = s->layer(layer_name);
        if (!l) return null_array();
        return view_array(st->neuron_data(l, key));
    } catch (...) {
        return null_array();
    }
}

ARRAY get_layer_data(STATE, char *, char *, char *) { return null_array(); }

ARRAY get_connection_data(STATE, char *, char *) { return null_array(); }

ARRAY get_weight_matrix(STATE state, char *conn_name, char *key) {
    if (!state || !conn_name) return null_array();
    try {
        State *st = static_cast<StateHandle *>(state)->state.get();
        Connection *c = st->network()->connection(conn_name);
        if (!c) return null_array();
        return view_array(st->weight_matrix(c, key ? key : "weights"));
    } catch (...) {
        return null_array();
    }
}

PROPS run(NETWORK net, ENVIRONMENT env, STATE state, PROPS args) {
    if (net == nullptr || state == nullptr) return nullptr;
    try {
        Engine e(static_cast<NetHandle *>(net)->net.get(),
                 env ? &static_cast<EnvHandle *>(env)->props : nullptr,
                 static_cast<StateHandle *>(state)->state.get());
        return e.run(args ? *P(args) : Props());
    } catch (const std::exception &ex) {
        if (!g_suppress) fprintf(stderr, "syngen run failed: %s\n", ex.what());
        return nullptr;
    } catch (...) {
        return nullptr;
    }
}

void destroy(void *obj) {
    if (!obj) return;
    if (!is_handle(obj)) {
        delete static_cast<Props *>(obj);
        return;
    }
    switch (*static_cast<uint32_t *>(obj)) {
        case TAG_NET: untrack_delete(static_cast<NetHandle *>(obj)); return;
        case TAG_ENV: untrack_delete(static_cast<EnvHandle *>(obj)); return;
        case TAG_STATE: untrack_delete(static_cast<StateHandle *>(obj)); return;
    }
}

void destroy_network(NETWORK net) {
    untrack_delete(static_cast<NetHandle *>(net));
}
void destroy_environment(ENVIRONMENT env) {
    untrack_delete(static_cast<EnvHandle *>(env));
}
void destroy_state(STATE state) {
    untrack_delete(static_cast<StateHandle *>(state));
}
void destroy_properties(PROPS props) { delete P(props); }

int get_num_gpus(void) {
    int n = sg_hip_device_count();
    return n < 0 ? 0 : n;
}

int get_cpu(void) { return get_num_gpus(); }

ARRAY get_gpus(void) {
    int n = get_num_gpus();
    int *p = (int *)malloc(sizeof(int) * (n > 0 ? n : 1));
    for (int i = 0; i < n; ++i) p[i] = i;
    return ARRAY{n, INT_POINTER, p, true};
}

ARRAY get_all_devices(void) {
    int n = get_num_gpus();
    int *p = (int *)malloc(sizeof(int) * (n + 1));
    for (int i = 0; i <= n; ++i) p[i] = i;
    return ARRAY{n + 1, INT_POINTER, p, true};
}

void set_suppress_output(bool val) {
    g_suppress = val;
    sg::set_host_messages(g_warnings, g_suppress);
}
void set_warnings(bool val) {
    g_warnings = val;
    sg::set_host_messages(g_warnings, g_suppress);
}
void set_debug(bool val) { g_debug = val; }

void interrupt_engine(void) { Engine::interrupt(); }

void add_io_callback(char *name, long long addr) {
    auto &c = Callbacks::get();
    c.add(c.io, name, (IoCallback)addr);
}
void add_weight_callback(char *name, long long addr) {
    auto &c = Callbacks::get();
    c.add(c.weight, name, (WeightCallback)addr);
}
void add_indices_weight_callback(char *name, long long addr) {
    auto &c = Callbacks::get();
    c.add(c.indices, name, (IndicesCallback)addr);
}
void add_distance_weight_callback(char *name, long long addr) {
    auto &c = Callbacks::get();
    c.add(c.distance, name, (DistanceCallback)addr);
}
void add_delay_weight_callback(char *name, long long addr) {
    auto &c = Callbacks::get();
    c.add(c.delay, name, (DelayCallback)addr);
}

void initialize_mpi(void) {}
int get_mpi_rank(void) { return 0; }
int get_mpi_size(void) { return 1; }
void finalize_mpi(void) {}

}  // extern "C"
