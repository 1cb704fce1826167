#include "engine.h"

#include <hip/hip_runtime.h>

#include <chrono>
#include <deque>
#include <functional>
#include <cstring>
#include <map>
#include <mutex>
#include <set>

#include "callbacks.h"

namespace sg {

namespace {

std::mutex g_engine_lock;
bool g_running = false;
std::atomic<bool> g_interrupt(false);

enum { IO_INPUT = 1, IO_OUTPUT = 2 };

// Pinned host staging buffer of one (layer, key).
struct HostBuf {
    float *p = nullptr;
    int n = 0;
    bool dirty = false;
    ~HostBuf() { if (p) (void)hipHostFree(p); }
    void alloc(int size) {
        n = size;
        check_hip(hipHostMalloc((void **)&p, (size_t)size * 4,
                                hipHostMallocDefault), "hipHostMalloc");
        std::memset(p, 0, (size_t)size * 4);
    }
};

struct ModuleLayer {
    LayerState *ls;
    int io = 0;
    std::set<std::string> in_keys, out_keys;
    IoCallback fn = nullptr;
    int id = 0;
};

// The "callback" module (io/impl/callback_module.cpp, io/module.cpp:84-110).
struct Module {
    std::vector<ModuleLayer> layers;
    bool clear_first = false;
    int start_delay = 0, cutoff = 0, rate = 1, curr = 0;
};

// Optional per-kernel timing with HIP events recorded on the engine's
// stream around each launch ("kernel timing": "true" in the run args).
struct KernelTimer {
    bool on = false;
    hipStream_t stream = nullptr;
    struct Rec { std::string name; hipEvent_t a, b; };
    std::vector<Rec> recs;
    template <class F>
    void time(const char *name, F &&launch) {
        if (!on) { launch(); return; }
        Rec r{name, nullptr, nullptr};
        check_hip(hipEventCreate(&r.a), "hipEventCreate");
        check_hip(hipEventCreate(&r.b), "hipEventCreate");
        check_hip(hipEventRecord(r.a, stream), "hipEventRecord");
        launch();
        check_hip(hipEventRecord(r.b, stream), "hipEventRecord");
        recs.push_back(r);
    }
    void report(Props &out) {
        std::map<std::string, std::pair<double, long>> acc;
        for (auto &r : recs) {
            float ms = 0.0f;
            check_hip(hipEventElapsedTime(&ms, r.a, r.b), "hipEventElapsedTime");
            acc[r.name].first += ms;
            acc[r.name].second += 1;
        }
        for (auto &kv : acc) {
            out.set(kv.first + " ms", std::to_string(kv.second.first));
            out.set(kv.first + " launches", std::to_string(kv.second.second));
        }
    }
    ~KernelTimer() {
        for (auto &r : recs) {
            (void)hipEventDestroy(r.a);
            (void)hipEventDestroy(r.b);
        }
    }
};

// Layer order of a sequential or feedforward cluster
// (engine/cluster/sequential_cluster.cpp:16-77): a breadth-first search
// backwards from the structure's output layers (no non-recurrent connection
// into the structure, network/layer.cpp:54-60) or layers without outgoing
// connections; a layer is placed once every layer it feeds in the structure
// (not itself, not through recurrent connections) is placed, and it goes in
// front of them.  Where the reference would requeue forever (a cycle of
// non-recurrent connections) this throws.
std::vector<Layer *> sequential_order(const Structure &st) {
    auto same = [&](const Layer *l) { return l->structure == &st; };
    auto is_output = [&](const Layer *l) {
        for (auto *c : l->outputs)
            if (same(c->to) && !c->recurrent) return false;
        return true;
    };
    std::deque<Layer *> queue;
    for (auto &l : st.layers)
        if (is_output(l.get()) || l->outputs.empty()) queue.push_back(l.get());
    std::set<Layer *> visited;
    std::vector<Layer *> order;
    size_t stalled = 0;
    while (!queue.empty()) {
        Layer *curr = queue.front();
        queue.pop_front();
        if (visited.count(curr)) continue;
        bool later = false;
        for (auto *c : curr->outputs)
            if (c->to != curr && same(c->to) && !c->recurrent &&
                !visited.count(c->to))
                later = true;
        if (later) {
            queue.push_back(curr);
            if (++stalled > 2 * queue.size() + st.layers.size())
                throw Error("Sequential cluster " + st.name + ": the layers "
                            "form a cycle of non-recurrent connections");
            continue;
        }
        stalled = 0;
        visited.insert(curr);
        order.insert(order.begin(), curr);
        for (auto *c : curr->inputs)
            if (c->from != curr && same(c->from) && !c->recurrent &&
                !visited.count(c->from))
                queue.push_back(c->from);
    }
    if (visited.size() < st.layers.size())
        throw Error("Sequential cluster failed to process all layers!");
    return order;
}

// Feedforward clusters refuse cycles reachable from a structure input layer
// (feedforward_cluster.cpp:8-38; every same-structure connection counts,
// self-connections included).
void check_feedforward(const Structure &st) {
    std::set<const Layer *> path;
    std::function<bool(const Layer *)> dfs = [&](const Layer *l) {
        if (path.count(l)) return false;
        path.insert(l);
        for (auto *c : l->outputs)
            if (c->to->structure == &st && !dfs(c->to)) return false;
        path.erase(l);
        return true;
    };
    for (auto &l : st.layers) {
        bool input = true;      // network/layer.cpp:46-52
        for (auto *c : l->inputs)
            if (c->from->structure == &st && !c->recurrent) input = false;
        if (input && !dfs(l.get()))
            throw Error("Feedforward engine requires a structure with no "
                        "cycles!");
    }
}

void check_layer_model_supports(LayerState *ls, const std::string &key) {
    if (key == "input" || key == "output") return;
    if (!ls->vars.count(key))
        throw Error("Failed to retrieve neuron data \"" + key + "\"");
}

}  // namespace

void Engine::interrupt() {
    std::lock_guard<std::mutex> g(g_engine_lock);
    if (g_running) g_interrupt = true;
}

Props *Engine::run(const Props &args) {
    {
        std::lock_guard<std::mutex> g(g_engine_lock);
        if (g_running) throw Error("Cannot run more than one engine at once!");
        g_running = true;
        g_interrupt = false;
    }
    struct Release {
        ~Release() {
            std::lock_guard<std::mutex> g(g_engine_lock);
            g_running = false;
            g_interrupt = false;
        }
    } release;

    // Devices: GPUs are 0..n-1, the host is n (resource_manager.h:53).
    int ngpu = sg_hip_device_count();
    if (ngpu <= 0)
        throw Error("No HIP device available: this library runs the "
                    "simulation on an MI355X GPU");
    std::vector<std::string> devs;
    if (args.has("devices")) devs.push_back(args.get("devices"));
    for (auto &d : args.get_array("devices")) devs.push_back(d);
    int device = 0;
    for (auto &d : devs) {
        int id = std::stoi(d);
        if (id >= ngpu)
            throw Error("Device " + d + " is not a GPU: this library "
                        "simulates on GPUs only");
        device = id;
    }
    check_hip(hipSetDevice(device), "hipSetDevice");

    // Cluster orders (engine/cluster/*.cpp): parallel structures activate
    // every layer, then update every layer; sequential and feedforward
    // structures activate and update layer by layer in dependency order
    // (sequential_cluster.cpp:100-113), weights backwards.
    std::vector<std::vector<Layer *>> seq_order(net_->structures().size());
    std::set<const Layer *> in_seq;
    for (size_t i = 0; i < net_->structures().size(); ++i) {
        const Structure &st = *net_->structures()[i];
        if (st.cluster == CLUSTER_PARALLEL) continue;
        if (st.cluster == CLUSTER_FEEDFORWARD) check_feedforward(st);
        seq_order[i] = sequential_order(st);
        for (auto &l : st.layers) in_seq.insert(l.get());
    }

    // "compact records": "false" keeps every per-synapse field in the
    // Izhikevich records uploaded by this run.
    state_->to_device(args.get_bool("compact records", true));

    const bool learning = args.get_bool("learning flag", true);
    const bool suppress = args.get_bool("suppress output", false);
    const int env_rate = std::max(1, args.get_int("environment rate", 1));
    long iterations = args.has("iterations") ? args.get_int("iterations", 1)
                                             : 0;

    // ---- environment: modules and buffers --------------------------------
    std::vector<Module> modules;
    std::map<LayerState *, int> io_types;
    std::map<LayerState *, std::set<std::string>> in_keys, out_keys;
    if (env_) {
        for (auto *mc : env_->get_child_array("modules")) {
            if (mc->get_bool("skip", false)) continue;
            const std::string type = mc->get("type", "");
            if (type != "callback")
                throw Error("Unrecognized module: " + type + "!");
            Module m;
            m.clear_first = mc->get_bool("clear", false);
            m.start_delay = mc->get_int("delay", 0);
            m.cutoff = mc->get_int("cutoff", 0);
            m.rate = mc->get_int("rate", 1);
            if (m.start_delay < 0 || m.cutoff < 0 || m.rate < 1)
                throw Error("Invalid module delay / cutoff / rate");
            for (auto *lc : mc->get_child_array("layers")) {
                Layer *l = lc->has("structure")
                    ? (net_->structure(lc->get("structure"))
                           ? net_->structure(lc->get("structure"))
                                 ->layer(lc->get("layer"))
                           : nullptr)
                    : net_->layer(lc->get("layer"));
                if (!l) throw Error("Could not find layer " + lc->get("layer"));
                ModuleLayer ml;
                ml.ls = state_->layer_state(l);
                if (lc->get_bool("input", false)) ml.io |= IO_INPUT;
                if (lc->get_bool("output", false)) ml.io |= IO_OUTPUT;
                for (auto &k : lc->get_array("input keys")) {
                    ml.in_keys.insert(k);
                    ml.io |= IO_INPUT;
                }
                for (auto &k : lc->get_array("output keys")) {
                    ml.out_keys.insert(k);
                    ml.io |= IO_OUTPUT;
                }
                if (ml.io == 0)
                    throw Error("Module has zero IO type for layer " + l->name);
                if (ml.in_keys.empty() && (ml.io & IO_INPUT))
                    ml.in_keys.insert("input");
                if (ml.out_keys.empty() && (ml.io & IO_OUTPUT))
                    ml.out_keys.insert("output");
                for (auto &k : ml.in_keys) check_layer_model_supports(ml.ls, k);
                for (auto &k : ml.out_keys) check_layer_model_supports(ml.ls, k);
                if (!lc->has("function"))
                    throw Error("Unspecified callback function for layer " +
                                l->name + " in CallbackModule!");
                if (!lc->has("id"))
                    throw Error("Unspecified callback id for layer " +
                                l->name + " in CallbackModule!");
                auto &cb = Callbacks::get();
                ml.fn = cb.find(cb.io, lc->get("function"), "io");
                ml.id = lc->get_int("id", 0);
                io_types[ml.ls] |= ml.io;
                for (auto &k : ml.in_keys) in_keys[ml.ls].insert(k);
                for (auto &k : ml.out_keys) out_keys[ml.ls].insert(k);
                m.layers.push_back(ml);
            }
            if (!m.layers.empty()) modules.push_back(m);
        }
    }
    for (auto &kv : io_types) {
        if ((kv.second & IO_INPUT) && in_keys[kv.first].empty())
            in_keys[kv.first].insert("input");
        if ((kv.second & IO_OUTPUT) && out_keys[kv.first].empty())
            out_keys[kv.first].insert("output");
    }
    auto is_input = [&](LayerState *ls) {
        auto it = io_types.find(ls);
        return it != io_types.end() && (it->second & IO_INPUT);
    };

    std::map<std::pair<LayerState *, std::string>, std::unique_ptr<HostBuf>>
        in_bufs, out_bufs;
    std::map<LayerState *, float *> inter;   // last fed input, on device
    std::vector<void *> dev_allocs;
    for (auto &kv : in_keys)
        for (auto &k : kv.second) {
            in_bufs[{kv.first, k}].reset(new HostBuf());
            in_bufs[{kv.first, k}]->alloc(kv.first->layer->size);
            if (k == "input") {
                float *d = nullptr;
                check_hip(hipMalloc((void **)&d,
                                    (size_t)kv.first->layer->size * 4),
                          "hipMalloc");
                check_hip(hipMemset(d, 0, (size_t)kv.first->layer->size * 4),
                          "hipMemset");
                inter[kv.first] = d;
                dev_allocs.push_back(d);
            }
        }
    for (auto &kv : out_keys)
        for (auto &k : kv.second) {
            out_bufs[{kv.first, k}].reset(new HostBuf());
            out_bufs[{kv.first, k}]->alloc(kv.first->layer->size);
        }
    struct FreeAll {
        std::vector<void *> &v;
        ~FreeAll() { for (auto *p : v) (void)hipFree(p); }
    } free_all{dev_allocs};

    hipStream_t stream;
    check_hip(hipStreamCreate(&stream), "hipStreamCreate");
    struct StreamGuard {
        hipStream_t s;
        ~StreamGuard() { (void)hipStreamDestroy(s); }
    } sg_{stream};
    // A second stream for the base activators of non-input layers, which
    // then run beside the conductance passes (see the main loop).
    hipStream_t side;
    check_hip(hipStreamCreate(&side), "hipStreamCreate");
    StreamGuard sg2_{side};
    // The fork and join events order two streams of one device: no
    // system-scope fence (that is for the host and other devices).
    hipEvent_t ev_fork, ev_join;
    const unsigned ev_flags = hipEventDisableTiming | hipEventDisableSystemFence;
    check_hip(hipEventCreateWithFlags(&ev_fork, ev_flags), "hipEventCreate");
    check_hip(hipEventCreateWithFlags(&ev_join, ev_flags), "hipEventCreate");
    struct EventGuard {
        hipEvent_t a, b;
        ~EventGuard() { (void)hipEventDestroy(a); (void)hipEventDestroy(b); }
    } eg_{ev_fork, ev_join};
    KernelTimer kt;
    kt.on = args.get_bool("kernel timing", false);
    kt.stream = stream;

    // ---- work folded into the Izhikevich neuron launch ---------------------
    // (filled before the main loop, read by the helpers below)
    struct InitFold { int noise = 0, dist = 3, op = SG_ADD;
                      float a = 0.0f, b = 0.0f, val = 0.0f; };
    std::map<LayerState *, InitFold> init_fold;     // init value drawn there
    std::set<ConnState *> store_sum;                // raw-sum activators
    std::map<LayerState *, std::vector<ConnState *>> tss_out;
    std::set<ConnState *> tss_folded;               // stepped there

    // ---- per-connection launch helpers -----------------------------------
    auto conn_desc = [&](ConnState &cs) {
        Connection &c = *cs.conn;
        sg_conn d;
        std::memset(&d, 0, sizeof(d));
        d.type = c.type;
        d.opcode = c.opcode;
        d.out_type = cs.from->out_type;
        d.delay = c.delay;
        d.convolutional = c.convolutional;
        d.from_size = c.from->size; d.from_rows = c.from->rows;
        d.from_cols = c.from->cols;
        d.to_size = c.to->size; d.to_rows = c.to->rows; d.to_cols = c.to->cols;
        d.row_field = c.arbor.row_field; d.col_field = c.arbor.col_field;
        d.row_stride = c.arbor.row_stride; d.col_stride = c.arbor.col_stride;
        d.row_spacing = c.arbor.row_spacing;
        d.col_spacing = c.arbor.col_spacing;
        d.row_offset = c.arbor.row_offset; d.col_offset = c.arbor.col_offset;
        d.wrap = c.arbor.wrap;
        d.from_row_start = c.subset.from_row_start;
        d.from_row_end = c.subset.from_row_end;
        d.from_col_start = c.subset.from_col_start;
        d.from_col_end = c.subset.from_col_end;
        d.to_row_start = c.subset.to_row_start;
        d.to_row_end = c.subset.to_row_end;
        d.to_col_start = c.subset.to_col_start;
        d.to_col_end = c.subset.to_col_end;
        d.k_per_dest = cs.k_per_dest;
        d.num_dest = cs.num_dest;
        d.num_weights = c.num_weights;
        const int per_word = cs.from->out_type == SG_OUT_BIT ? 32 : 1;
        d.src = cs.from->output.du() +
                (size_t)(c.delay / per_word) * c.from->size;
        d.dst = cs.to->input.df() + (size_t)c.reg * c.to->size;
        d.weights = cs.arr("weights").df();
        // Plastic base-activator weights stay finite: the updater clamps
        // them to [1e-4, max weight].
        d.skip_silent = d.out_type == SG_OUT_BIT && cs.finite_weights;
        // Delay-0 reads of an Izhikevich layer use the plane its neuron
        // kernel wrote last step (packed once at run start); other BIT
        // sources are packed per activation into the connection's own.
        if (cs.from->model == Model::IZHIKEVICH && c.delay == 0 &&
            cs.from->d_plane) {
            d.plane = cs.from->d_plane;
            d.plane_ready = 1;
        } else {
            d.plane = cs.d_plane;
        }
        d.store_sum = store_sum.count(&cs) ? 1 : 0;
        d.sparse_from = cs.d_sparse_from;
        d.tile_counts = cs.d_tile_counts;
        // A second-order host activates with its slaves' copy of its
        // weights (synapse_data.cpp:20-23), which may hold any value.
        if (c.so_host) {
            d.weights = cs.arr("second order weights").df();
            d.skip_silent = 0;
        }
        return d;
    };
    auto iz_syn = [&](ConnState &cs) {
        sg_iz_syn s;
        std::memset(&s, 0, sizeof(s));
        s.record = cs.packed ? cs.d_record : nullptr;
        s.weights = cs.arr("weights").df();
        s.short_traces = cs.arr("short trace").df();
        s.long_traces = cs.arr("long trace").df();
        s.presyn_traces = cs.arr("presyn trace").df();
        s.stp_r = cs.arr("short term R").df();
        s.stp_w = cs.arr("short term w").df();
        s.weight_derivatives = cs.arr("weight derivatives").df();
        s.delays = cs.d_delays;
        s.tss_cur = cs.d_tss[cs.tss_cur];
        s.src_stp = cs.compact ? cs.d_src_stp : nullptr;
        s.baseline_conductance = cs.baseline_conductance;
        s.learning_rate = cs.learning_rate;
        s.max_weight = cs.conn->max_weight;
        s.stp_flag = cs.stp_flag;
        return s;
    };
    auto iz_dest = [&](ConnState &cs, int kind) {
        sg_iz_dest d;
        std::memset(&d, 0, sizeof(d));
        LayerState &t = *cs.to;
        d.out_word0 = t.output.du();
        d.post_trace = t.var("post trace").df();
        d.time_since_spike = t.var("time since spike").di();
        d.voltage = t.var("voltage").df();
        d.input = t.input.df();
        // ACTIV_EXTRACTIONS_SHORT / _LONG taus are floats
        // (izhikevich_attributes.cpp:219-249).
        switch (kind) {
            case SG_IZ_ADD:
                d.short_cond = t.var("ampa").df(); d.short_tau = (float)0.8;
                d.long_cond = t.var("nmda").df(); d.long_tau = (float)0.993;
                break;
            case SG_IZ_SUB:
                d.short_cond = t.var("gabaa").df(); d.short_tau = (float)0.833;
                d.long_cond = t.var("gabab").df(); d.long_tau = (float)0.993;
                break;
            case SG_IZ_MULT:
                d.short_cond = t.var("mult").df(); d.short_tau = (float)0.95;
                break;
            case SG_IZ_REWARD:
                d.short_cond = t.var("dopamine").df(); d.short_tau = (float)0.95;
                break;
        }
        return d;
    };
    auto fusable = [&](ConnState &cs) {
        return cs.iz && !cs.direct && cs.iz_kind >= 0 &&
               cs.iz_kind != SG_IZ_GAP && cs.update_kind >= 0;
    };
    auto activate = [&](ConnState &cs, hipStream_t st) {
        if (cs.fused_gol) return;
        sg_conn d = conn_desc(cs);
        d.beside = st == side;
        if (!cs.iz || cs.direct) {
            kt.time("activate_base", [&] {
                check_hip(sg_hip_activate_base(&d, st), "activate");
            });
            return;
        }
        sg_iz_dest dd = iz_dest(cs, cs.iz_kind);
        sg_iz_syn s = iz_syn(cs);
        if (cs.iz_kind == SG_IZ_GAP) {
            kt.time("iz_gap", [&] {
                check_hip(sg_hip_iz_activate(&d, &s, &dd, SG_IZ_GAP, st),
                          "gap activate");
            });
            return;
        }
        if (!tss_folded.count(&cs)) kt.time("tss_step", [&] {
            check_hip(sg_hip_tss_step(cs.d_tss[1 - cs.tss_cur],
                                      cs.d_tss[cs.tss_cur], cs.d_visits,
                                      cs.d_rev_delays, d.src,
                                      cs.conn->from->size, st),
                      "tss step");
        });
        if (cs.pending_update) {
            kt.time("iz_update_activate", [&] {
                check_hip(sg_hip_iz_update_activate(
                              &d, &s, &dd, cs.iz_kind, cs.update_kind,
                              state_->d_src_table(), state_->d_dst_table(),
                              st),
                          "update+activate");
            });
            cs.pending_update = false;
        } else {
            kt.time("iz_activate", [&] {
                check_hip(sg_hip_iz_activate(&d, &s, &dd, cs.iz_kind, st),
                          "iz activate");
            });
        }
        if (!tss_folded.count(&cs)) {
            cs.tss_cur = 1 - cs.tss_cur;
            // A compact record's per-source state takes this step's
            // updater step here (the neuron kernel does it for folded
            // connections).
            if (cs.compact && learning && cs.update_kind >= 0)
                check_hip(sg_hip_iz_source_update(
                              cs.d_src_stp, cs.d_tss[cs.tss_cur],
                              cs.conn->from->size, cs.update_kind == 1,
                              cs.stp_flag, st),
                          "source update");
        }
    };
    auto update = [&](ConnState &cs) {
        sg_conn d = conn_desc(cs);
        sg_iz_syn s = iz_syn(cs);
        sg_iz_dest dd = iz_dest(cs, cs.iz_kind >= 0 ? cs.iz_kind : SG_IZ_ADD);
        kt.time("iz_update", [&] {
            check_hip(sg_hip_iz_update(&d, &s, &dd, cs.update_kind,
                                       state_->d_src_table(),
                                       state_->d_dst_table(), stream),
                      "iz update");
        });
    };

    // ---- game-of-life fast path -------------------------------------------
    // A game-of-life layer whose only input is one 3x3 convolutional
    // self-connection (stride 1, offsets -1, add, delay 0), with one history
    // word and no environment I/O, is stepped by sg_hip_gol_step.
    std::map<LayerState *, ConnState *> gol_fast;
    std::map<LayerState *, uint32_t *> gol_spare, gol_home;
    for (auto &lsp : state_->layers()) {
        LayerState *ls = lsp.get();
        if (ls->model != Model::GAME_OF_LIFE || ls->history != 1) continue;
        if (io_types.count(ls) || ls->layer->config.has_child("init config") ||
            in_seq.count(ls->layer))
            continue;
        if (ls->layer->inputs.size() != 1 || ls->layer->registers > 1)
            continue;
        Connection *c = ls->layer->inputs[0];
        const ArborConfig &a = c->arbor;
        if (c->from != ls->layer || c->type != SG_CONVERGENT ||
            !c->convolutional || c->opcode != SG_ADD || c->delay != 0 ||
            a.row_field != 3 || a.col_field != 3 || a.row_stride != 1 ||
            a.col_stride != 1 || a.row_spacing != 1 || a.col_spacing != 1 ||
            a.row_offset != -1 || a.col_offset != -1)
            continue;
        ConnState *cs = state_->conn_state(c);
        cs->fused_gol = true;
        gol_fast[ls] = cs;
        uint32_t *spare = nullptr;
        check_hip(hipMalloc((void **)&spare, ls->output.bytes()), "hipMalloc");
        dev_allocs.push_back(spare);
        gol_spare[ls] = spare;
        gol_home[ls] = ls->output.du();
    }
    struct GolReset {
        std::map<LayerState *, ConnState *> &m;
        ~GolReset() { for (auto &kv : m) kv.second->fused_gol = false; }
    } gol_reset{gol_fast};

    // ---- init configs (instruction.cpp:3-30) ------------------------------
    auto init_layer = [&](LayerState *ls, bool overwrite_default, long step) {
        Props *ic = ls->layer->config.get_child("init config");
        const int n = ls->layer->size;
        float *in = ls->input.df();
        if (!ic) {
            if (overwrite_default)
                check_hip(sg_hip_set_f32(in, 0.0f, n, 1, stream), "clear");
            return;
        }
        const bool ow = overwrite_default;
        const std::string type = ic->get("type", "normal");
        const uint64_t seed = ls->layer->id;
        if (type == "flat") {
            check_hip(sg_hip_set_f32(in, ic->get_float("value", 1.0f), n, ow,
                                     stream), "init flat");
        } else if (type == "uniform") {
            check_hip(sg_hip_init_random(in, n, 0, ic->get_float("min", 0.0f),
                                         ic->get_float("max", 1.0f), 0, ow,
                                         seed, step, stream), "init uniform");
        } else if (type == "normal") {
            check_hip(sg_hip_init_random(in, n, 1,
                                         ic->get_float("mean", 1.0f),
                                         ic->get_float("std dev", 0.1f), 0, ow,
                                         seed, step, stream), "init normal");
        } else if (type == "poisson") {
            const float rate = (float)(0.001 * ic->get_float("rate", 1.0f));
            check_hip(sg_hip_init_random(in, n, 2, rate, 0.0f,
                                         ic->get_float("value", 20.0f), ow,
                                         seed, step, stream), "init poisson");
        } else {
            throw Error("Unrecognized init type: " + type);
        }
    };

    // ---- folding plans ----------------------------------------------------
    // Init configs (instruction.cpp:3-30) of non-input Izhikevich layers:
    // the value -- 0 for a plain clear -- is drawn by the neuron kernel when
    // no base activator writes the input (it is then the input), or when
    // exactly one does and covers every neuron (it stores its raw sum and
    // the neuron kernel applies the activator's own aggregate to the value
    // and the sum, the reference's operation).  Gap junctions add to the
    // input too, so their layers keep the separate pass.
    for (auto &lsp : state_->layers()) {
        LayerState *ls = lsp.get();
        if (ls->model != Model::IZHIKEVICH || io_types.count(ls) ||
            ls->layer->registers > 1)
            continue;
        InitFold f;
        if (Props *ic = ls->layer->config.get_child("init config")) {
            const std::string type = ic->get("type", "normal");
            if (type == "flat") {
                f.dist = 3; f.a = ic->get_float("value", 1.0f);
            } else if (type == "uniform") {
                f.dist = 0; f.a = ic->get_float("min", 0.0f);
                f.b = ic->get_float("max", 1.0f);
            } else if (type == "normal") {
                f.dist = 1; f.a = ic->get_float("mean", 1.0f);
                f.b = ic->get_float("std dev", 0.1f);
            } else if (type == "poisson") {
                f.dist = 2;
                f.a = (float)(0.001 * ic->get_float("rate", 1.0f));
                f.val = ic->get_float("value", 20.0f);
            } else {
                continue;    // refused by init_layer with its message
            }
        }
        std::vector<ConnState *> base;
        bool gap = false;
        for (auto *c : ls->layer->inputs) {
            ConnState *cs = state_->conn_state(c);
            if (!cs->iz || cs->direct) base.push_back(cs);
            else if (cs->iz_kind == SG_IZ_GAP) gap = true;
        }
        if (gap) continue;
        if (base.empty()) {
            f.noise = 1;
        } else if (base.size() == 1 && base[0]->conn->type != SG_SUBSET) {
            f.noise = 2;
            f.op = base[0]->conn->opcode;
            store_sum.insert(base[0]);
        } else {
            continue;
        }
        init_fold[ls] = f;
    }
    // Time since spike of conductance connections from Izhikevich layers is
    // stepped by the source layer's neuron kernel (same rule as
    // sg_hip_tss_step, on the same pre-shift words, after the activation
    // that reads the previous value).
    for (auto &csp : state_->conns()) {
        ConnState *cs = csp.get();
        if (!cs->iz || cs->direct || cs->iz_kind < 0 ||
            cs->iz_kind == SG_IZ_GAP || cs->from->model != Model::IZHIKEVICH)
            continue;
        // In a sequential structure the source may update before this
        // connection activates within a step; step it at activation then.
        if (in_seq.count(cs->from->layer) || in_seq.count(cs->to->layer))
            continue;
        auto &v = tss_out[cs->from];
        if ((int)v.size() >= SG_MAX_TSS_OUT) continue;
        v.push_back(cs);
        tss_folded.insert(cs);
    }
    for (auto &lsp : state_->layers())
        if (lsp->d_plane)
            check_hip(sg_hip_pack_plane(lsp->output.du(), lsp->layer->size, 0,
                                        lsp->d_plane, stream),
                      "pack plane");

    // Activation order of each layer: its dendritic tree depth first --
    // every internal node's subtree, then that node's register folded into
    // its parent's (cluster_node.cpp:115-212, kernel.cpp:161-187) -- and
    // then the node's own connections; without dendrites, the connections
    // in creation order.  Internal registers start each run at their
    // node's init value.
    // kinds: a connection's activation, a register fold, a second-order
    // node's weight copy, a slave's fold into it, the host's activation
    enum { ACT_CONN, ACT_FOLD, ACT_SO_COPY, ACT_SO_SLAVE, ACT_SO_HOST };
    struct Act {
        int kind;
        ConnState *cs, *host;
        const Dendrite *child, *parent;
    };
    std::function<bool(const ConnState &)> is_base_act;
    std::map<LayerState *, std::vector<Act>> plan;
    bool any_dendrites = false;
    for (auto &lsp : state_->layers()) {
        LayerState *ls = lsp.get();
        std::vector<Act> &v = plan[ls];
        std::function<void(const Dendrite *)> dfs = [&](const Dendrite *d) {
            for (const Dendrite *ch : d->children) {
                dfs(ch);
                v.push_back(Act{ACT_FOLD, nullptr, nullptr, ch, d});
            }
            if (d->second_order) {
                // cluster_node.cpp:118-121, 185-211: copy the host's
                // weights, let every slave fold into the copy, then
                // activate the host with it
                if (!d->host) return;
                ConnState *host = state_->conn_state(d->host);
                v.push_back(Act{ACT_SO_COPY, host, host, nullptr, nullptr});
                for (Connection *c : d->conns)
                    v.push_back(Act{ACT_SO_SLAVE, state_->conn_state(c), host,
                                    nullptr, nullptr});
                v.push_back(Act{ACT_SO_HOST, host, host, nullptr, nullptr});
                return;
            }
            for (Connection *c : d->conns)
                v.push_back(Act{ACT_CONN, state_->conn_state(c), nullptr,
                                nullptr, nullptr});
        };
        dfs(ls->layer->dendrites[0].get());
        const int n = ls->layer->size;
        for (size_t i = 1; i < ls->layer->dendrites.size(); ++i) {
            any_dendrites = true;
            const Dendrite &d = *ls->layer->dendrites[i];
            check_hip(sg_hip_set_f32(ls->input.df() + (size_t)d.reg * n,
                                     d.init, n, 1, stream),
                      "dendrite init");
        }
    }
    auto activate_layer = [&](LayerState *ls, bool split_base) {
        const int n = ls->layer->size;
        for (const Act &a : plan[ls]) {
            if (a.kind == ACT_CONN || a.kind == ACT_SO_HOST) {
                activate(*a.cs, split_base && is_base_act(*a.cs) ? side
                                                                : stream);
                continue;
            }
            if (a.kind == ACT_SO_COPY) {
                const Array &w = a.host->arr("weights");
                check_hip(hipMemcpyAsync(a.host->arr("second order weights").dev,
                                         w.dev, w.dev_elems * 4,
                                         hipMemcpyDeviceToDevice, stream),
                          "second order copy");
                continue;
            }
            if (a.kind == ACT_SO_SLAVE) {
                const sg_conn d = conn_desc(*a.cs);
                check_hip(sg_hip_activate_second_order(
                              &d, a.host->arr("second order weights").df(),
                              stream),
                          "second order");
                continue;
            }
            float *in = ls->input.df();
            check_hip(sg_hip_dendrite(in + (size_t)a.parent->reg * n,
                                      in + (size_t)a.child->reg * n, n,
                                      a.child->opcode, a.child->init,
                                      stream),
                      "dendrite");
        }
    };

    // Split the pre-input activators over two streams only when both kinds
    // are there (and no dendritic tree orders a layer's activators).  The
    // activators that aggregate into a layer's input -- base activators and
    // gap junctions -- all go to the second stream, in order.
    auto is_base = [&](const ConnState &cs) {
        return (!cs.iz || cs.direct || cs.iz_kind == SG_IZ_GAP) &&
               !cs.fused_gol;
    };
    is_base_act = is_base;
    bool any_base = false, any_other = false;
    for (auto &lsp : state_->layers()) {
        if (is_input(lsp.get()) || gol_fast.count(lsp.get()) ||
            in_seq.count(lsp->layer))
            continue;
        for (auto *c : lsp->layer->inputs) {
            const ConnState &cs = *state_->conn_state(c);
            if (cs.fused_gol) continue;
            (is_base(cs) ? any_base : any_other) = true;
        }
    }
    // "concurrent activators": "false" keeps every activator on one stream.
    const bool split_ok = any_base && any_other && !any_dendrites &&
                          args.get_bool("concurrent activators", true);

    // ---- main loop --------------------------------------------------------
    auto &layers = state_->layers();
    long done = 0;
    const auto t0 = std::chrono::steady_clock::now();
    bool interrupted = false;
    for (long it = 0; iterations == 0 || it < iterations; ++it) {
        // 1. pre-input: init values, then every activator.  Base activators
        // (latency-bound) go to the second stream and run beside the
        // conductance passes (HBM-bound): they write the layers' inputs, the
        // conductance passes write conductances and synapse records, and
        // both read only last step's spikes; the join below precedes the
        // neuron launch.  With kernel timing on, one stream.
        for (auto &lsp : layers) {
            LayerState *ls = lsp.get();
            if (is_input(ls) || gol_fast.count(ls) || in_seq.count(ls->layer))
                continue;
            if (!init_fold.count(ls)) init_layer(ls, true, it);
        }
        const bool split = split_ok && !kt.on;
        if (split) {
            check_hip(hipEventRecord(ev_fork, stream), "hipEventRecord");
            check_hip(hipStreamWaitEvent(side, ev_fork, 0), "wait fork");
        }
        for (auto &lsp : layers) {
            LayerState *ls = lsp.get();
            if (is_input(ls) || gol_fast.count(ls) || in_seq.count(ls->layer))
                continue;
            activate_layer(ls, split);
        }
        if (split) {
            check_hip(hipEventRecord(ev_join, side), "hipEventRecord");
            check_hip(hipStreamWaitEvent(stream, ev_join, 0), "wait join");
        }
        // 2. environment input
        bool any_input = false;
        for (auto &m : modules) {
            if (m.start_delay > 0) continue;
            for (auto &ml : m.layers) {
                if (!(ml.io & IO_INPUT)) continue;
                if (m.cutoff == 0 || m.curr < m.cutoff) {
                    if (m.curr % m.rate != 0) continue;
                    for (auto &k : ml.in_keys) {
                        if (!any_input) {
                            check_hip(hipStreamSynchronize(stream), "sync");
                            any_input = true;
                        }
                        HostBuf &b = *in_bufs.at({ml.ls, k});
                        if (m.clear_first) std::memset(b.p, 0, (size_t)b.n * 4);
                        ml.fn(ml.id, ml.ls->layer->size, b.p);
                        b.dirty = true;
                    }
                } else if (m.curr == m.cutoff) {
                    if (!any_input) {
                        check_hip(hipStreamSynchronize(stream), "sync");
                        any_input = true;
                    }
                    HostBuf &b = *in_bufs.at({ml.ls, "input"});
                    std::memset(b.p, 0, (size_t)b.n * 4);
                    b.dirty = true;
                }
            }
        }
        for (auto &kv : in_keys) {
            LayerState *ls = kv.first;
            const size_t bytes = (size_t)ls->layer->size * 4;
            for (auto &k : kv.second) {
                HostBuf &b = *in_bufs.at({ls, k});
                if (k == "input") {
                    if (b.dirty)
                        check_hip(hipMemcpyAsync(inter[ls], b.p, bytes,
                                                 hipMemcpyHostToDevice, stream),
                                  "input H2D");
                    b.dirty = false;
                    check_hip(hipMemcpyAsync(ls->input.dev, inter[ls], bytes,
                                             hipMemcpyDeviceToDevice, stream),
                              "input D2D");
                } else {
                    check_hip(hipMemcpyAsync(ls->var(k).dev, b.p, bytes,
                                             hipMemcpyHostToDevice, stream),
                              "aux input H2D");
                }
            }
        }
        // 3. post-input, per structure in order: the input layers of a
        // parallel structure activate; a sequential structure runs its
        // layers one by one (init, activators, neuron update)
        // (parallel_cluster.cpp:112-120, sequential_cluster.cpp:100-108).
        const bool last = iterations != 0 && it + 1 == iterations;
        auto update_neurons = [&](const std::vector<LayerState *> &lss) {
            std::vector<sg_iz_layer> iz_layers;
            std::vector<ConnState *> stepped;   // time since spike stepped
            for (LayerState *ls : lss) {
                const int n = ls->layer->size;
                switch (ls->model) {
                    case Model::IZHIKEVICH: {
                        sg_iz_layer l;
                        std::memset(&l, 0, sizeof(l));
                        l.size = n; l.history = ls->history;
                        l.input = ls->input.df(); l.output = ls->output.du();
                        l.ampa = ls->var("ampa").df(); l.nmda = ls->var("nmda").df();
                        l.gabaa = ls->var("gabaa").df();
                        l.gabab = ls->var("gabab").df();
                        l.mult = ls->var("mult").df();
                        l.dopamine = ls->var("dopamine").df();
                        l.voltage = ls->var("voltage").df();
                        l.recovery = ls->var("recovery").df();
                        l.post_trace = ls->var("post trace").df();
                        l.time_since_spike = ls->var("time since spike").di();
                        l.a = ls->var("a").df(); l.b = ls->var("b").df();
                        l.c = ls->var("c").df(); l.d = ls->var("d").df();
                        l.plane = ls->d_plane;
                        for (ConnState *cs : tss_out[ls]) {
                            stepped.push_back(cs);
                            sg_tss_out &o = l.tss[l.ntss++];
                            o.next = cs->d_tss[1 - cs->tss_cur];
                            o.cur = cs->d_tss[cs->tss_cur];
                            o.visits = cs->d_visits;
                            o.rev_delays = cs->d_rev_delays;
                            o.word = cs->conn->delay / 32;
                            o.stp = (cs->compact && learning &&
                                     cs->update_kind >= 0)
                                ? cs->d_src_stp : nullptr;
                            o.stp_sub = cs->update_kind == 1;
                            o.stp_flag = cs->stp_flag;
                        }
                        auto fi = init_fold.find(ls);
                        if (fi != init_fold.end()) {
                            const InitFold &f = fi->second;
                            l.noise = f.noise; l.noise_dist = f.dist;
                            l.noise_op = f.op;
                            l.noise_a = f.a; l.noise_b = f.b; l.noise_val = f.val;
                            l.noise_seed = ls->layer->id;
                            l.noise_step = (uint64_t)it;
                        }
                        iz_layers.push_back(l);
                        break;
                    }
                    case Model::GAME_OF_LIFE: {
                        auto gf = gol_fast.find(ls);
                        if (gf != gol_fast.end()) {
                            uint32_t *spare = gol_spare[ls];
                            float k9[9];
                            std::memcpy(k9, gf->second->arr("weights").f(), 36);
                            kt.time("gol_step", [&] {
                                check_hip(sg_hip_gol_step(
                                              ls->output.du(), spare,
                                              last ? ls->input.df() : nullptr,
                                              ls->layer->rows, ls->layer->cols,
                                              gf->second->conn->arbor.wrap, k9,
                                              ls->survival_min, ls->survival_max,
                                              ls->birth_min, ls->birth_max,
                                              stream),
                                          "gol step");
                            });
                            gol_spare[ls] = ls->output.du();
                            ls->output.dev = spare;
                        } else {
                            check_hip(sg_hip_gol_neuron(
                                          ls->input.df(), ls->output.du(), n,
                                          ls->history, ls->survival_min,
                                          ls->survival_max, ls->birth_min,
                                          ls->birth_max, stream),
                                      "gol neuron");
                        }
                        break;
                    }
                    case Model::RATE_ENCODING:
                        check_hip(sg_hip_rate_neuron(ls->input.df(),
                                                     ls->output.df(), n,
                                                     ls->history, stream),
                                  "rate neuron");
                        break;
                    case Model::RELAY:
                        check_hip(sg_hip_relay_neuron(ls->input.df(),
                                                      ls->output.df(), n,
                                                      ls->history, ls->ramp,
                                                      stream),
                                  "relay neuron");
                        break;
                }
            }
            if (!iz_layers.empty()) {
                kt.time("iz_neuron", [&] {
                    check_hip(sg_hip_iz_neurons(iz_layers.data(),
                                                (int)iz_layers.size(), stream),
                              "iz neuron");
                });
                for (ConnState *cs : stepped) cs->tss_cur = 1 - cs->tss_cur;
            }
        };
        for (size_t si = 0; si < net_->structures().size(); ++si) {
            const Structure &st = *net_->structures()[si];
            if (st.cluster == CLUSTER_PARALLEL) {
                for (auto &l : st.layers) {
                    LayerState *ls = state_->layer_state(l.get());
                    if (!is_input(ls)) continue;
                    // Only an "input" key makes the layer's input register
                    // the environment's (cluster_node.cpp:20-44); auxiliary
                    // input keys leave it to the usual clear.
                    const auto ik = in_keys.find(ls);
                    init_layer(ls, ik == in_keys.end() ||
                                   !ik->second.count("input"), it);
                    activate_layer(ls, false);
                }
                continue;
            }
            for (Layer *l : seq_order[si]) {
                LayerState *ls = state_->layer_state(l);
                if (is_input(ls)) {
                    const auto ik = in_keys.find(ls);
                    init_layer(ls, ik == in_keys.end() ||
                                   !ik->second.count("input"), it);
                } else if (!init_fold.count(ls)) {
                    init_layer(ls, true, it);
                }
                activate_layer(ls, false);
                update_neurons({ls});
            }
        }
        // 4. neuron updates of the parallel structures (every Izhikevich
        // layer in one launch)
        {
            std::vector<LayerState *> par;
            for (auto &lsp : layers)
                if (!in_seq.count(lsp->layer)) par.push_back(lsp.get());
            update_neurons(par);
        }
        // 5. weight updates
        if (learning) {
            for (auto &csp : state_->conns()) {
                ConnState &cs = *csp;
                if (cs.update_kind < 0) continue;
                if (fusable(cs)) cs.pending_update = true;
                else update(cs);
            }
        }
        // 6. output
        if (it % env_rate == 0 && !suppress && !out_keys.empty()) {
            for (auto &kv : out_keys) {
                LayerState *ls = kv.first;
                const size_t bytes = (size_t)ls->layer->size * 4;
                for (auto &k : kv.second) {
                    HostBuf &b = *out_bufs.at({ls, k});
                    const void *src = k == "output" ? ls->output.dev
                                    : k == "input" ? ls->input.dev
                                    : ls->var(k).dev;
                    check_hip(hipMemcpyAsync(b.p, src, bytes,
                                             hipMemcpyDeviceToHost, stream),
                              "output D2H");
                }
            }
            check_hip(hipStreamSynchronize(stream), "output sync");
            for (auto &m : modules) {
                if (m.start_delay > 0) continue;
                if (!(m.cutoff == 0 || m.curr < m.cutoff)) continue;
                if (m.curr % m.rate != 0) continue;
                for (auto &ml : m.layers) {
                    if (!(ml.io & IO_OUTPUT)) continue;
                    for (auto &k : ml.out_keys)
                        ml.fn(ml.id, ml.ls->layer->size,
                              out_bufs.at({ml.ls, k})->p);
                }
            }
        }
        // modules cycle (io/module.cpp:98-108)
        for (auto &m : modules) {
            if (m.start_delay <= 0) {
                if (m.cutoff == 0 || m.curr < m.cutoff) ++m.curr;
                else if (m.cutoff != 0 && m.curr == m.cutoff) ++m.curr;
            } else {
                --m.start_delay;
            }
        }
        check_hip(hipGetLastError(), "step");
        ++done;
        if (g_interrupt) {
            interrupted = true;
            break;
        }
    }
    // Apply updates still deferred into a next activation.
    for (auto &csp : state_->conns())
        if (csp->pending_update) {
            update(*csp);
            csp->pending_update = false;
        }
    // Leave double-buffered game-of-life words in the state's own buffer.
    for (auto &kv : gol_fast) {
        LayerState *ls = kv.first;
        if (ls->output.dev != gol_home[ls]) {
            check_hip(hipMemcpyAsync(gol_home[ls], ls->output.dev,
                                     ls->output.bytes(),
                                     hipMemcpyDeviceToDevice, stream),
                      "gol restore");
            ls->output.dev = gol_home[ls];
        }
    }
    check_hip(hipStreamSynchronize(stream), "final sync");
    Props timing;
    kt.report(timing);
    const double secs = std::chrono::duration<double>(
        std::chrono::steady_clock::now() - t0).count();

    // Report (report.cpp:8-16, engine.cpp:584-590).
    Props *r = new Props();
    r->set("iterations", std::to_string(done));
    r->set("total time", std::to_string((float)secs));
    r->set("average time", std::to_string((float)(secs / (double)done)));
    r->set("fps", std::to_string((float)((double)done / secs)));
    r->set("network bytes", std::to_string(state_->network_bytes()));
    size_t buf_bytes = 0;
    for (auto &kv : in_bufs) buf_bytes += (size_t)kv.second->n * 4;
    for (auto &kv : out_bufs) buf_bytes += (size_t)kv.second->n * 4;
    r->set("state buffer bytes", std::to_string(inter.size() ? buf_bytes : 0));
    r->set("engine buffer bytes", std::to_string(buf_bytes));
    r->set_child("args", args);
    r->set("interrupted", interrupted ? "true" : "false");
    // 2: the input-writing activators ran on a second stream beside the
    // conductance passes; 1: every activator on one stream.
    r->set("activator streams", split_ok && !kt.on ? "2" : "1");
    // Izhikevich connections whose records are compact (sg_hip.h SG_IZC_*).
    int compact = 0;
    for (auto &csp : state_->conns()) compact += csp->compact ? 1 : 0;
    r->set("compact records", std::to_string(compact));
    Props mem;
    mem.set("device", std::to_string(device));
    mem.set("bytes", std::to_string(state_->network_bytes()));
    r->add_to_child_array("memory usage", mem);
    if (kt.on) r->set_child("kernel timing", timing);
    return r;
}

}  // namespace sg
