#include "engine.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <deque>
#include <functional>
#include <initializer_list>
#include <map>
#include <mutex>
#include <set>
#include <unordered_map>

#include "callbacks.h"

namespace sg {

namespace {

std::mutex g_engine_lock;
bool g_running = false;
std::atomic<bool> g_interrupt(false);

enum { IO_INPUT = 1, IO_OUTPUT = 2 };

// ---------------------------------------------------------------------------
// Scheduler: streams and the dependencies between their launches.
//
// Every launch (a "task") names the arrays it reads and writes.  A task
// waits for the last writer of everything it touches and, for what it
// writes, for every reader since that write -- but only where those ran on
// another stream; launches on one stream are ordered by the stream.  When
// more than one stream is in use, each task records an event from its
// stream's ring, and a wait targets the event of the exact task it depends
// on (or a later one on the same stream, if the ring has come round: that
// only waits longer).  This keeps the serial order's results while letting
// independent launches overlap.  Inside a HIP graph capture the same
// records and waits become the graph's edges.
// ---------------------------------------------------------------------------

enum ResKind { RES_IN = 0, RES_OUT = 1, RES_STATE = 2, RES_SYN = 3,
               RES_TSS = 4, RES_CTR = 5 };
typedef uint64_t Res;

Res res(ResKind k, const void *p) {
    return ((uint64_t)(uintptr_t)p << 3) | (uint64_t)k;
}

class Sched {
  public:
    ~Sched() {
        for (auto &s : st_) {
            if (s.s) (void)hipStreamSynchronize(s.s);
            for (auto e : s.ring) (void)hipEventDestroy(e);
            if (s.s) (void)hipStreamDestroy(s.s);
        }
    }

    // At least n streams (created once, kept between runs; all at the
    // default priority).
    void ensure(int n) {
        const unsigned flags = hipEventDisableTiming |
                               hipEventDisableSystemFence;
        while ((int)st_.size() < n) {
            Str s;
            check_hip(hipStreamCreate(&s.s), "hipStreamCreate");
            s.ring.resize(kRing);
            s.ring_seq.assign(kRing, 0);
            for (auto &e : s.ring)
                check_hip(hipEventCreateWithFlags(&e, flags),
                          "hipEventCreate");
            st_.push_back(s);
        }
        for (auto &s : st_) s.seen.resize(st_.size(), 0);
    }
    hipStream_t stream(int i) const { return st_[i].s; }
    int count() const { return (int)st_.size(); }

    // Streams that take part in this run: 1 keeps everything on stream 0
    // and records no events.
    void set_used(int n) {
        used_ = n;
        res_.clear();
    }
    bool multi() const { return used_ > 1; }

    hipStream_t begin(int i, std::initializer_list<Res> reads,
                      std::initializer_list<Res> writes) {
        cur_r_.assign(reads.begin(), reads.end());
        cur_w_.assign(writes.begin(), writes.end());
        return start(i);
    }
    hipStream_t begin(int i, const std::vector<Res> &reads,
                      const std::vector<Res> &writes) {
        cur_r_ = reads;
        cur_w_ = writes;
        return start(i);
    }

    void end() {
        const int i = cur_;
        const uint64_t seq = ++st_[i].seq;
        if (multi()) {
            mark(i);
            for (Res r : cur_r_) {
                Entry &e = entry(r);
                e.readers[i] = seq;
            }
            for (Res w : cur_w_) {
                Entry &e = entry(w);
                e.writer = i;
                e.wseq = seq;
                std::fill(e.readers.begin(), e.readers.end(), 0);
            }
        }
        cur_ = -1;
    }

  private:
    hipStream_t start(int i) {
        if (!multi()) i = 0;
        cur_ = i;
        if (multi()) {
            for (Res r : cur_r_) {
                auto it = res_.find(r);
                if (it != res_.end()) dep(i, it->second.writer,
                                          it->second.wseq);
            }
            for (Res w : cur_w_) {
                auto it = res_.find(w);
                if (it == res_.end()) continue;
                dep(i, it->second.writer, it->second.wseq);
                for (int j = 0; j < used_; ++j)
                    dep(i, j, it->second.readers[j]);
            }
        }
        return st_[i].s;
    }

  public:
    // Stream 0 waits for every other stream's work so far.
    void join() {
        if (!multi()) return;
        for (int j = 1; j < used_; ++j) {
            ++st_[j].seq;
            mark(j);
            dep(0, j, st_[j].seq);
        }
    }
    // Every other stream waits for stream 0's work so far.
    void fork() {
        if (!multi()) return;
        ++st_[0].seq;
        mark(0);
        for (int j = 1; j < used_; ++j) dep(j, 0, st_[0].seq);
    }
    // Forget the recorded readers and writers: only valid once the streams
    // are ordered by a join and a fork.
    void reset() { res_.clear(); }
    void barrier() {
        join();
        fork();
        reset();
    }
    void synchronize() {
        for (int j = 0; j < used_ && j < (int)st_.size(); ++j)
            check_hip(hipStreamSynchronize(st_[j].s), "stream sync");
    }

  private:
    static constexpr int kRing = 64;
    struct Str {
        hipStream_t s = nullptr;
        uint64_t seq = 0;
        std::vector<hipEvent_t> ring;
        std::vector<uint64_t> ring_seq;   // task whose record a slot holds
        std::vector<uint64_t> seen;       // per stream: already waited for
    };
    struct Entry {
        int writer = -1;
        uint64_t wseq = 0;
        std::vector<uint64_t> readers;
    };
    Entry &entry(Res r) {
        Entry &e = res_[r];
        if ((int)e.readers.size() < used_) e.readers.resize(used_, 0);
        return e;
    }
    void mark(int i) {
        Str &s = st_[i];
        const int slot = (int)(s.seq % kRing);
        check_hip(hipEventRecord(s.ring[slot], s.s), "hipEventRecord");
        s.ring_seq[slot] = s.seq;
    }
    void dep(int i, int j, uint64_t q) {
        if (j < 0 || j == i || q == 0) return;
        Str &s = st_[i];
        if (s.seen[j] >= q) return;
        Str &o = st_[j];
        const int slot = (int)(q % kRing);
        check_hip(hipStreamWaitEvent(s.s, o.ring[slot], 0),
                  "hipStreamWaitEvent");
        s.seen[j] = std::max(s.seen[j], o.ring_seq[slot]);
    }

    std::vector<Str> st_;
    int used_ = 1;
    std::unordered_map<Res, Entry> res_;
    int cur_ = -1;
    std::vector<Res> cur_r_, cur_w_;
};

// A captured run of steps and the host state it was captured from.
struct StepGraph {
    std::vector<uint64_t> key;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    ~StepGraph() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
    }
};

struct Cache : EngineCache {
    Sched sched;
    // second word buffer of each game-of-life layer on the fused step
    std::map<LayerState *, uint32_t *> gol_spare;
    // step counter that captured steps read their random draws' step from
    uint64_t *d_step = nullptr;
    uint64_t *h_step = nullptr;     // pinned staging for it
    std::vector<std::unique_ptr<StepGraph>> graphs;
    ~Cache() override {
        graphs.clear();
        for (auto &kv : gol_spare) (void)hipFree(kv.second);
        if (d_step) (void)hipFree(d_step);
        if (h_step) (void)hipHostFree(h_step);
    }
};

Cache &cache_of(State &s) {
    if (!s.engine_cache) s.engine_cache.reset(new Cache());
    return static_cast<Cache &>(*s.engine_cache);
}

// Pinned host staging buffer of one (layer, key), mapped into the device's
// address space and coherent (the device reads and writes it over the
// link, past its caches), and the event of the last copy that used it.
struct HostBuf {
    float *p = nullptr;
    void *dp = nullptr;    // the device's address of p
    int n = 0;
    bool dirty = false;
    hipEvent_t done = nullptr;
    bool pending = false;
    ~HostBuf() {
        if (done) (void)hipEventDestroy(done);
        if (p) (void)hipHostFree(p);
    }
    void alloc(int size) {
        n = size;
        check_hip(hipHostMalloc((void **)&p, (size_t)size * 4,
                                hipHostMallocMapped | hipHostMallocCoherent),
                  "hipHostMalloc");
        std::memset(p, 0, (size_t)size * 4);
        check_hip(hipHostGetDevicePointer(&dp, p, 0),
                  "hipHostGetDevicePointer");
        check_hip(hipEventCreateWithFlags(&done, hipEventDisableTiming),
                  "hipEventCreate");
    }
    // the copy that used the buffer last has finished
    void wait() {
        if (!pending) return;
        check_hip(hipEventSynchronize(done), "hipEventSynchronize");
        pending = false;
    }
    void record(hipStream_t s) {
        check_hip(hipEventRecord(done, s), "hipEventRecord");
        pending = true;
    }
};

// Input staging of one (layer, key): two pinned buffers, so that the host
// fills one while a copy from the other may still be in flight.
struct InBuf {
    HostBuf b[2];
    int cur = 0;        // the buffer written last
    bool dirty = false; // written since its last copy to the device
    void alloc(int size) { b[0].alloc(size); b[1].alloc(size); }
    // The buffer the host writes next (its previous copy has finished);
    // it starts as a copy of the last input, so a callback that writes
    // part of it sees the rest as the single buffer would hold it.
    float *next(bool clear) {
        HostBuf &x = b[1 - cur];
        x.wait();
        if (clear) std::memset(x.p, 0, (size_t)x.n * 4);
        else std::memcpy(x.p, b[cur].p, (size_t)x.n * 4);
        cur = 1 - cur;
        dirty = true;
        return x.p;
    }
    HostBuf &last() { return b[cur]; }
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

// Optional per-kernel timing with HIP events recorded on stream 0 around
// each launch ("kernel timing": "true" in the run args; every launch then
// runs on stream 0).
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

// ---- kernel descriptors (engine/kernel/synapse_data.cpp:9-64) -------------

sg_conn conn_desc(ConnState &cs, bool store_sum) {
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
    d.store_sum = store_sum ? 1 : 0;
    d.sparse_from = cs.d_sparse_from;
    d.tile_counts = cs.d_tile_counts;
    // A second-order host activates with its slaves' copy of its
    // weights (synapse_data.cpp:20-23), which may hold any value.
    if (c.so_host) {
        d.weights = cs.arr("second order weights").df();
        d.skip_silent = 0;
    }
    return d;
}

sg_iz_syn iz_syn(ConnState &cs) {
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
}

sg_iz_dest iz_dest(ConnState &cs, int kind) {
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
}

void launch_update(State &state, ConnState &cs, hipStream_t st) {
    sg_conn d = conn_desc(cs, false);
    sg_iz_syn s = iz_syn(cs);
    sg_iz_dest dd = iz_dest(cs, cs.iz_kind >= 0 ? cs.iz_kind : SG_IZ_ADD);
    check_hip(sg_hip_iz_update(&d, &s, &dd, cs.update_kind,
                               state.d_src_table(), state.d_dst_table(), st),
              "iz update");
}

bool fusable(const ConnState &cs) {
    return cs.iz && !cs.direct && cs.iz_kind >= 0 &&
           cs.iz_kind != SG_IZ_GAP && cs.update_kind >= 0;
}

}  // namespace

void flush_pending_updates(State &state) {
    bool any = false;
    for (auto &csp : state.conns())
        if (csp->pending_update) {
            if (!any) check_hip(hipDeviceSynchronize(), "flush sync");
            any = true;
            launch_update(state, *csp, nullptr);
            csp->pending_update = false;
        }
    if (any) check_hip(hipDeviceSynchronize(), "flush");
}

void Engine::interrupt() {
    std::lock_guard<std::mutex> g(g_engine_lock);
    if (g_running) g_interrupt = true;
}

namespace {

// One run of the engine over a state: the reference's single-thread loop
// (engine/engine.cpp:226-315) with its parallel, sequential and feedforward
// clusters (engine/cluster/*.cpp).
class Runner {
  public:
    Runner(Network *net, const Props *env, State *state, const Props &args)
        : net_(net), env_(env), st_(state), args_(args) {}
    Props *run();

  private:
    // What one step's random draws and last-step work depend on.
    struct StepCtx {
        uint64_t step;            // the draws' step (added to *base if set)
        const uint64_t *base;     // device step counter of captured steps
        bool last;                // the run's last step
    };
    struct InitFold { int noise = 0, dist = 3, op = SG_ADD;
                      float a = 0.0f, b = 0.0f, val = 0.0f; };
    // a layer's activation plan: a connection's activation, a register
    // fold, a second-order node's weight copy, a slave's fold into it, the
    // host's activation
    enum { ACT_CONN, ACT_FOLD, ACT_SO_COPY, ACT_SO_SLAVE, ACT_SO_HOST };
    struct Act {
        int kind;
        ConnState *cs, *host;
        const Dendrite *child, *parent;
    };

    Sched &sched() { return C_->sched; }
    bool is_input(LayerState *ls) const {
        auto it = io_types_.find(ls);
        return it != io_types_.end() && (it->second & IO_INPUT);
    }
    bool any_input_key() const {
        for (auto &kv : in_keys_)
            if (kv.second.count("input")) return true;
        return false;
    }
    bool is_base(const ConnState &cs) const {
        return (!cs.iz || cs.direct || cs.iz_kind == SG_IZ_GAP) &&
               !cs.fused_gol;
    }
    int main_of(const LayerState *ls) const {
        return main_.at(ls->layer->structure);
    }
    int side_of(const LayerState *ls) const {
        return side_.at(ls->layer->structure);
    }
    bool split_of(const LayerState *ls) const {
        return side_of(ls) != main_of(ls);
    }
    uint64_t global(long it) const { return step0_ + (uint64_t)it; }

    void setup_modules();
    void setup_plans();
    void setup_streams();

    void init_layer(LayerState *ls, bool overwrite_default, const StepCtx &x);
    void activate(ConnState &cs, int sidx);
    void activate_layer(LayerState *ls, bool split);
    void update_neurons(const std::vector<LayerState *> &lss, int sidx,
                        const StepCtx &x);

    void pre_input(const StepCtx &x);
    void input_phase();
    void post_input(const StepCtx &x);
    void neurons(const StepCtx &x);
    void updates();
    bool output_issue(long it);
    void output_deliver();
    void cycle_modules();

    std::vector<uint64_t> graph_key(int block) const;
    StepGraph *step_graph(int block);
    StepGraph *cached_graph(int block) const;

    Network *net_;
    const Props *env_;
    State *st_;
    const Props &args_;
    Cache *C_ = nullptr;

    bool learning_ = true, suppress_ = false;
    int env_rate_ = 1;
    long iterations_ = 0;
    uint64_t step0_ = 0;

    std::vector<std::vector<Layer *>> seq_order_;
    std::set<const Layer *> in_seq_;

    std::vector<Module> modules_;
    std::map<LayerState *, int> io_types_;
    std::map<LayerState *, std::set<std::string>> in_keys_, out_keys_;
    std::map<std::pair<LayerState *, std::string>, std::unique_ptr<InBuf>>
        in_bufs_;
    std::map<std::pair<LayerState *, std::string>, std::unique_ptr<HostBuf>>
        out_bufs_;

    KernelTimer kt_;

    std::map<LayerState *, InitFold> init_fold_;
    std::set<ConnState *> store_sum_;
    std::map<LayerState *, std::vector<ConnState *>> tss_out_;
    std::set<ConnState *> tss_folded_;
    std::map<LayerState *, ConnState *> gol_fast_;
    std::map<LayerState *, uint32_t *> gol_home_;
    // game-of-life layers that nothing else reads (their only outgoing
    // connection is the self-connection): stepped up to
    // sg_hip_gol_max_steps steps per launch in eager runs; gol_ahead_ counts the steps an earlier launch already
    // took beyond the current one, gol_rem_ the steps left in the run from
    // the current one (1 while steps are captured into graphs)
    std::set<LayerState *> gol_multi_;
    std::map<LayerState *, int> gol_ahead_;
    long gol_rem_ = 1;
    long gol_launches_ = 0;
    bool gol_pending() const {
        for (const auto &kv : gol_ahead_)
            if (kv.second > 0) return true;
        return false;
    }
    std::map<LayerState *, std::vector<Act>> plan_;
    bool any_dendrites_ = false, any_second_order_ = false;

    std::map<const Structure *, int> main_, side_;
    int nstreams_ = 1;
    bool any_split_ = false;
};

// ---- environment: modules and buffers ---------------------------------------
void Runner::setup_modules() {
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
                ml.ls = st_->layer_state(l);
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
                io_types_[ml.ls] |= ml.io;
                for (auto &k : ml.in_keys) in_keys_[ml.ls].insert(k);
                for (auto &k : ml.out_keys) out_keys_[ml.ls].insert(k);
                m.layers.push_back(ml);
            }
            if (!m.layers.empty()) modules_.push_back(m);
        }
    }
    for (auto &kv : io_types_) {
        if ((kv.second & IO_INPUT) && in_keys_[kv.first].empty())
            in_keys_[kv.first].insert("input");
        if ((kv.second & IO_OUTPUT) && out_keys_[kv.first].empty())
            out_keys_[kv.first].insert("output");
    }
    for (auto &kv : in_keys_)
        for (auto &k : kv.second) {
            const int n = kv.first->layer->size;
            in_bufs_[{kv.first, k}].reset(new InBuf());
            in_bufs_[{kv.first, k}]->alloc(n);
        }
    for (auto &kv : out_keys_)
        for (auto &k : kv.second) {
            out_bufs_[{kv.first, k}].reset(new HostBuf());
            out_bufs_[{kv.first, k}]->alloc(kv.first->layer->size);
        }
}

// ---- plans ------------------------------------------------------------------
void Runner::setup_plans() {
    // Game-of-life fast path: a game-of-life layer whose only input is one
    // 3x3 convolutional self-connection (stride 1, offsets -1, add, delay
    // 0), with one history word and no environment I/O, is stepped by
    // sg_hip_gol_step over double-buffered words.
    for (auto &lsp : st_->layers()) {
        LayerState *ls = lsp.get();
        if (ls->model != Model::GAME_OF_LIFE || ls->history != 1) continue;
        if (io_types_.count(ls) ||
            ls->layer->config.has_child("init config") ||
            in_seq_.count(ls->layer))
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
        ConnState *cs = st_->conn_state(c);
        cs->fused_gol = true;
        gol_fast_[ls] = cs;
        if (ls->layer->outputs.size() == 1 && ls->layer->outputs[0] == c)
            gol_multi_.insert(ls);
        gol_ahead_[ls] = 0;
        if (!C_->gol_spare.count(ls)) {
            uint32_t *spare = nullptr;
            check_hip(hipMalloc((void **)&spare, ls->output.bytes()),
                      "hipMalloc");
            C_->gol_spare[ls] = spare;
        }
        gol_home_[ls] = ls->output.du();
    }

    // Init configs (instruction.cpp:3-30) of non-input Izhikevich layers:
    // the value -- 0 for a plain clear -- is drawn by the neuron kernel when
    // no base activator writes the input (it is then the input), or when
    // exactly one does and covers every neuron (it stores its raw sum and
    // the neuron kernel applies the activator's own aggregate to the value
    // and the sum, the reference's operation).  Gap junctions add to the
    // input too, so their layers keep the separate pass.
    for (auto &lsp : st_->layers()) {
        LayerState *ls = lsp.get();
        if (ls->model != Model::IZHIKEVICH || io_types_.count(ls) ||
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
            ConnState *cs = st_->conn_state(c);
            if (!cs->iz || cs->direct) base.push_back(cs);
            else if (cs->iz_kind == SG_IZ_GAP) gap = true;
        }
        if (gap) continue;
        if (base.empty()) {
            f.noise = 1;
        } else if (base.size() == 1 && base[0]->conn->type != SG_SUBSET) {
            f.noise = 2;
            f.op = base[0]->conn->opcode;
            store_sum_.insert(base[0]);
        } else {
            continue;
        }
        init_fold_[ls] = f;
    }
    // Time since spike of conductance connections from Izhikevich layers is
    // stepped by the source layer's neuron kernel (same rule as
    // sg_hip_tss_step, on the same pre-shift words, after the activation
    // that reads the previous value).
    for (auto &csp : st_->conns()) {
        ConnState *cs = csp.get();
        if (!cs->iz || cs->direct || cs->iz_kind < 0 ||
            cs->iz_kind == SG_IZ_GAP || cs->from->model != Model::IZHIKEVICH)
            continue;
        // In a sequential structure the source may update before this
        // connection activates within a step; step it at activation then.
        if (in_seq_.count(cs->from->layer) || in_seq_.count(cs->to->layer))
            continue;
        auto &v = tss_out_[cs->from];
        if ((int)v.size() >= SG_MAX_TSS_OUT) continue;
        v.push_back(cs);
        tss_folded_.insert(cs);
    }

    // Activation order of each layer: its dendritic tree depth first --
    // every internal node's subtree, then that node's register folded into
    // its parent's (cluster_node.cpp:115-212, kernel.cpp:161-187) -- and
    // then the node's own connections; without dendrites, the connections
    // in creation order.
    for (auto &lsp : st_->layers()) {
        LayerState *ls = lsp.get();
        std::vector<Act> &v = plan_[ls];
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
                any_second_order_ = true;
                ConnState *host = st_->conn_state(d->host);
                v.push_back(Act{ACT_SO_COPY, host, host, nullptr, nullptr});
                for (Connection *c : d->conns)
                    v.push_back(Act{ACT_SO_SLAVE, st_->conn_state(c), host,
                                    nullptr, nullptr});
                v.push_back(Act{ACT_SO_HOST, host, host, nullptr, nullptr});
                return;
            }
            for (Connection *c : d->conns)
                v.push_back(Act{ACT_CONN, st_->conn_state(c), nullptr,
                                nullptr, nullptr});
        };
        dfs(ls->layer->dendrites[0].get());
        if (ls->layer->dendrites.size() > 1) any_dendrites_ = true;
    }
}

// ---- streams ----------------------------------------------------------------
// One stream per structure (engine/cluster/cluster.cpp:9-16,
// parallel_cluster.cpp:27-40 give each layer of a cluster a compute stream
// and the cluster an IO stream; here a structure is the unit); in a
// parallel structure whose layers have both kinds of activators, the ones
// that write a layer's input (base activators, gap junctions) on a second
// stream of their own, beside the conductance passes (which write
// conductances and synapse records).  The environment's copies run on
// the stream of their layer.  Kernel timing keeps everything on stream 0.
void Runner::setup_streams() {
    const bool one = kt_.on;
    const bool per_structure = !one &&
                               args_.get_bool("structure streams", true);
    const bool concurrent = !one &&
                            args_.get_bool("concurrent activators", true);
    int next = 0;
    const auto &structs = net_->structures();
    for (size_t si = 0; si < structs.size(); ++si)
        main_[structs[si].get()] = per_structure ? next++ : 0;
    if (next == 0) next = 1;
    for (size_t si = 0; si < structs.size(); ++si) {
        const Structure *s = structs[si].get();
        bool any_base = false, any_other = false;
        if (s->cluster == CLUSTER_PARALLEL)
            for (auto &l : s->layers) {
                LayerState *ls = st_->layer_state(l.get());
                if (is_input(ls) || gol_fast_.count(ls)) continue;
                for (auto *c : l->inputs) {
                    const ConnState &cs = *st_->conn_state(c);
                    if (cs.fused_gol) continue;
                    (is_base(cs) ? any_base : any_other) = true;
                }
            }
        const bool split = concurrent && any_base && any_other &&
                           !any_dendrites_;
        side_[s] = split ? next++ : main_[s];
        any_split_ = any_split_ || split;
    }
    nstreams_ = next;
    sched().ensure(nstreams_);
    sched().set_used(nstreams_);
    kt_.stream = sched().stream(0);
}

// ---- launches ---------------------------------------------------------------
// init configs (instruction.cpp:3-30)
void Runner::init_layer(LayerState *ls, bool overwrite_default,
                        const StepCtx &x) {
    Props *ic = ls->layer->config.get_child("init config");
    const int n = ls->layer->size;
    float *in = ls->input.df();
    const int m = main_of(ls);
    const Res rin = res(RES_IN, ls);
    if (!ic) {
        if (!overwrite_default) return;
        hipStream_t s = sched().begin(m, {}, {rin});
        check_hip(sg_hip_set_f32(in, 0.0f, n, 1, s), "clear");
        sched().end();
        return;
    }
    const bool ow = overwrite_default;
    const std::string type = ic->get("type", "normal");
    const uint64_t seed = ls->layer->id;
    const Res ctr = res(RES_CTR, C_->d_step);
    int dist = -1;
    float a = 0.0f, b = 0.0f, val = 0.0f;
    if (type == "flat") {
        hipStream_t s = sched().begin(m, {rin}, {rin});
        check_hip(sg_hip_set_f32(in, ic->get_float("value", 1.0f), n, ow, s),
                  "init flat");
        sched().end();
        return;
    } else if (type == "uniform") {
        dist = 0; a = ic->get_float("min", 0.0f); b = ic->get_float("max", 1.0f);
    } else if (type == "normal") {
        dist = 1; a = ic->get_float("mean", 1.0f);
        b = ic->get_float("std dev", 0.1f);
    } else if (type == "poisson") {
        dist = 2; a = (float)(0.001 * ic->get_float("rate", 1.0f));
        val = ic->get_float("value", 20.0f);
    } else {
        throw Error("Unrecognized init type: " + type);
    }
    hipStream_t s = sched().begin(m, {rin, ctr}, {rin});
    check_hip(sg_hip_init_random(in, n, dist, a, b, val, ow, seed, x.step,
                                 x.base, s), "init random");
    sched().end();
}

void Runner::activate(ConnState &cs, int sidx) {
    if (cs.fused_gol) return;
    LayerState *from = cs.from, *to = cs.to;
    sg_conn d = conn_desc(cs, store_sum_.count(&cs) > 0);
    d.beside = sidx != main_of(to);
    const Res out_from = res(RES_OUT, from), syn = res(RES_SYN, &cs),
              tss = res(RES_TSS, &cs), in_to = res(RES_IN, to);
    if (!cs.iz || cs.direct) {
        // the connection's own plane scratch counts as its TSS resource
        hipStream_t s = sched().begin(sidx, {out_from, syn, in_to},
                                      {in_to, tss});
        kt_.time("activate_base", [&] {
            check_hip(sg_hip_activate_base(&d, s), "activate");
        });
        sched().end();
        return;
    }
    sg_iz_dest dd = iz_dest(cs, cs.iz_kind);
    sg_iz_syn sy = iz_syn(cs);
    if (cs.iz_kind == SG_IZ_GAP) {
        hipStream_t s = sched().begin(sidx, {res(RES_STATE, from), syn, in_to},
                                      {in_to});
        kt_.time("iz_gap", [&] {
            check_hip(sg_hip_iz_activate(&d, &sy, &dd, SG_IZ_GAP, s),
                      "gap activate");
        });
        sched().end();
        return;
    }
    const bool folded = tss_folded_.count(&cs) > 0;
    if (!folded) {
        hipStream_t s = sched().begin(sidx, {out_from, tss}, {tss});
        kt_.time("tss_step", [&] {
            check_hip(sg_hip_tss_step(cs.d_tss[1 - cs.tss_cur],
                                      cs.d_tss[cs.tss_cur], cs.d_visits,
                                      cs.d_rev_delays, d.src,
                                      cs.conn->from->size, s),
                      "tss step");
        });
        sched().end();
    }
    const Res state_to = res(RES_STATE, to), out_to = res(RES_OUT, to);
    hipStream_t s = sched().begin(sidx, {out_from, tss, syn, state_to, out_to},
                                  {syn, state_to});
    if (cs.pending_update) {
        kt_.time("iz_update_activate", [&] {
            check_hip(sg_hip_iz_update_activate(
                          &d, &sy, &dd, cs.iz_kind, cs.update_kind,
                          st_->d_src_table(), st_->d_dst_table(), s),
                      "update+activate");
        });
        cs.pending_update = false;
    } else {
        kt_.time("iz_activate", [&] {
            check_hip(sg_hip_iz_activate(&d, &sy, &dd, cs.iz_kind, s),
                      "iz activate");
        });
    }
    sched().end();
    if (!folded) {
        cs.tss_cur = 1 - cs.tss_cur;
        // A compact record's per-source state takes this step's updater
        // step here (the neuron kernel does it for folded connections).
        if (cs.compact && learning_ && cs.update_kind >= 0) {
            hipStream_t s2 = sched().begin(sidx, {tss}, {tss});
            check_hip(sg_hip_iz_source_update(
                          cs.d_src_stp, cs.d_tss[cs.tss_cur],
                          cs.conn->from->size, cs.update_kind == 1,
                          cs.stp_flag, s2),
                      "source update");
            sched().end();
        }
    }
}

void Runner::activate_layer(LayerState *ls, bool split) {
    const int n = ls->layer->size;
    const int m = main_of(ls), sd = side_of(ls);
    const Res rin = res(RES_IN, ls);
    for (const Act &a : plan_[ls]) {
        if (a.kind == ACT_CONN || a.kind == ACT_SO_HOST) {
            activate(*a.cs, split && is_base(*a.cs) ? sd : m);
            continue;
        }
        if (a.kind == ACT_SO_COPY) {
            const Res hs = res(RES_SYN, a.host);
            hipStream_t s = sched().begin(m, {hs}, {hs});
            const Array &w = a.host->arr("weights");
            check_hip(hipMemcpyAsync(a.host->arr("second order weights").dev,
                                     w.dev, w.dev_elems * 4,
                                     hipMemcpyDeviceToDevice, s),
                      "second order copy");
            sched().end();
            continue;
        }
        if (a.kind == ACT_SO_SLAVE) {
            const sg_conn d = conn_desc(*a.cs, false);
            const Res hs = res(RES_SYN, a.host);
            hipStream_t s = sched().begin(
                m, {res(RES_OUT, a.cs->from), res(RES_SYN, a.cs), hs}, {hs});
            check_hip(sg_hip_activate_second_order(
                          &d, a.host->arr("second order weights").df(), s),
                      "second order");
            sched().end();
            continue;
        }
        float *in = ls->input.df();
        hipStream_t s = sched().begin(m, {rin}, {rin});
        check_hip(sg_hip_dendrite(in + (size_t)a.parent->reg * n,
                                  in + (size_t)a.child->reg * n, n,
                                  a.child->opcode, a.child->init, s),
                  "dendrite");
        sched().end();
    }
}

void Runner::update_neurons(const std::vector<LayerState *> &lss, int sidx,
                            const StepCtx &x) {
    std::vector<sg_iz_layer> iz_layers;
    std::vector<Res> reads, writes;
    std::vector<ConnState *> stepped;   // time since spike stepped
    for (LayerState *ls : lss) {
        const int n = ls->layer->size;
        const Res rin = res(RES_IN, ls), rout = res(RES_OUT, ls),
                  rst = res(RES_STATE, ls);
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
                reads.insert(reads.end(), {rin, rout, rst});
                writes.insert(writes.end(), {rin, rout, rst});
                for (ConnState *cs : tss_out_[ls]) {
                    stepped.push_back(cs);
                    sg_tss_out &o = l.tss[l.ntss++];
                    o.next = cs->d_tss[1 - cs->tss_cur];
                    o.cur = cs->d_tss[cs->tss_cur];
                    o.visits = cs->d_visits;
                    o.rev_delays = cs->d_rev_delays;
                    o.word = cs->conn->delay / 32;
                    o.stp = (cs->compact && learning_ &&
                             cs->update_kind >= 0)
                        ? cs->d_src_stp : nullptr;
                    o.stp_sub = cs->update_kind == 1;
                    o.stp_flag = cs->stp_flag;
                    const Res t = res(RES_TSS, cs);
                    reads.push_back(t);
                    writes.push_back(t);
                }
                auto fi = init_fold_.find(ls);
                if (fi != init_fold_.end()) {
                    const InitFold &f = fi->second;
                    l.noise = f.noise; l.noise_dist = f.dist;
                    l.noise_op = f.op;
                    l.noise_a = f.a; l.noise_b = f.b; l.noise_val = f.val;
                    l.noise_seed = ls->layer->id;
                    l.noise_step = x.step;
                    l.noise_base = x.base;
                    if (x.base) reads.push_back(res(RES_CTR, C_->d_step));
                }
                iz_layers.push_back(l);
                break;
            }
            case Model::GAME_OF_LIFE: {
                auto gf = gol_fast_.find(ls);
                if (gf != gol_fast_.end()) {
                    // taken by an earlier launch of several steps
                    int &ahead = gol_ahead_[ls];
                    if (ahead > 0) {
                        --ahead;
                        break;
                    }
                    float k9[9];
                    std::memcpy(k9, gf->second->arr("weights").f(), 36);
                    const int k = gol_multi_.count(ls)
                        ? (int)std::min<long>(
                              sg_hip_gol_max_steps(k9, ls->layer->rows,
                                                   ls->layer->cols),
                              std::max<long>(1, gol_rem_))
                        : 1;
                    // the input (neighbour sums) is written at the run's
                    // last step only
                    const bool last = k == 1 ? x.last
                                             : iterations_ != 0 && k == gol_rem_;
                    uint32_t *spare = C_->gol_spare.at(ls);
                    hipStream_t s = sched().begin(
                        sidx, {rout, rin, res(RES_SYN, gf->second)},
                        {rout, rin});
                    kt_.time("gol_step", [&] {
                        check_hip(sg_hip_gol_steps(
                                      ls->output.du(), spare,
                                      last ? ls->input.df() : nullptr,
                                      ls->layer->rows, ls->layer->cols,
                                      gf->second->conn->arbor.wrap, k9,
                                      ls->survival_min, ls->survival_max,
                                      ls->birth_min, ls->birth_max, k, s),
                                  "gol step");
                    });
                    sched().end();
                    ahead = k - 1;
                    ++gol_launches_;
                    C_->gol_spare[ls] = ls->output.du();
                    ls->output.dev = spare;
                } else {
                    hipStream_t s = sched().begin(sidx, {rin, rout}, {rout});
                    check_hip(sg_hip_gol_neuron(
                                  ls->input.df(), ls->output.du(), n,
                                  ls->history, ls->survival_min,
                                  ls->survival_max, ls->birth_min,
                                  ls->birth_max, s),
                              "gol neuron");
                    sched().end();
                }
                break;
            }
            case Model::RATE_ENCODING: {
                hipStream_t s = sched().begin(sidx, {rin, rout}, {rout});
                check_hip(sg_hip_rate_neuron(ls->input.df(), ls->output.df(),
                                             n, ls->history, s),
                          "rate neuron");
                sched().end();
                break;
            }
            case Model::RELAY: {
                hipStream_t s = sched().begin(sidx, {rin, rout}, {rout});
                check_hip(sg_hip_relay_neuron(ls->input.df(), ls->output.df(),
                                              n, ls->history, ls->ramp, s),
                          "relay neuron");
                sched().end();
                break;
            }
        }
    }
    if (!iz_layers.empty()) {
        hipStream_t s = sched().begin(sidx, reads, writes);
        kt_.time("iz_neuron", [&] {
            check_hip(sg_hip_iz_neurons(iz_layers.data(),
                                        (int)iz_layers.size(), s),
                      "iz neuron");
        });
        sched().end();
        for (ConnState *cs : stepped) cs->tss_cur = 1 - cs->tss_cur;
    }
}

// ---- phases of a step ---------------------------------------------------------
// 1. layers without environment input: init value (or clear), then every
// activator; they read outputs of the previous step.
void Runner::pre_input(const StepCtx &x) {
    auto &layers = st_->layers();
    for (auto &lsp : layers) {
        LayerState *ls = lsp.get();
        if (is_input(ls) || gol_fast_.count(ls) || in_seq_.count(ls->layer))
            continue;
        if (!init_fold_.count(ls)) init_layer(ls, true, x);
    }
    for (auto &lsp : layers) {
        LayerState *ls = lsp.get();
        if (is_input(ls) || gol_fast_.count(ls) || in_seq_.count(ls->layer))
            continue;
        activate_layer(ls, split_of(ls));
    }
}

// 2. environment input: callbacks fill pinned buffers, which are copied to
// the input layers on their structures' streams.
void Runner::input_phase() {
    for (auto &m : modules_) {
        if (m.start_delay > 0) continue;
        for (auto &ml : m.layers) {
            if (!(ml.io & IO_INPUT)) continue;
            if (m.cutoff == 0 || m.curr < m.cutoff) {
                if (m.curr % m.rate != 0) continue;
                for (auto &k : ml.in_keys) {
                    float *p = in_bufs_.at({ml.ls, k})->next(m.clear_first);
                    ml.fn(ml.id, ml.ls->layer->size, p);
                }
            } else if (m.curr == m.cutoff) {
                in_bufs_.at({ml.ls, "input"})->next(true);
            }
        }
    }
    // Each step the last buffer written (its values stay the input while a
    // module skips steps) is copied by a kernel on the layer's stream,
    // reading the pinned buffer over the link: no DMA engine, no stream
    // and events of its own.
    for (auto &kv : in_keys_) {
        LayerState *ls = kv.first;
        const long n = ls->layer->size;
        for (auto &k : kv.second) {
            InBuf &b = *in_bufs_.at({ls, k});
            const bool in = k == "input";
            void *dst = in ? ls->input.dev : ls->var(k).dev;
            const Res r = in ? res(RES_IN, ls) : res(RES_STATE, ls);
            hipStream_t s = sched().begin(main_of(ls), {}, {r});
            check_hip(sg_hip_copy_u32(b.last().dp, dst, n, s),
                      in ? "input copy" : "aux input copy");
            sched().end();
            b.last().record(s);
            b.dirty = false;
        }
    }
}

// 3. post-input, per structure in order: the input layers of a parallel
// structure activate; a sequential structure runs its layers one by one
// (init, activators, neuron update) (parallel_cluster.cpp:112-120,
// sequential_cluster.cpp:100-108).
void Runner::post_input(const StepCtx &x) {
    const auto &structs = net_->structures();
    for (size_t si = 0; si < structs.size(); ++si) {
        const Structure &st = *structs[si];
        if (st.cluster == CLUSTER_PARALLEL) {
            for (auto &l : st.layers) {
                LayerState *ls = st_->layer_state(l.get());
                if (!is_input(ls)) continue;
                // Only an "input" key makes the layer's input register the
                // environment's (cluster_node.cpp:20-44); auxiliary input
                // keys leave it to the usual clear.
                const auto ik = in_keys_.find(ls);
                init_layer(ls, ik == in_keys_.end() ||
                               !ik->second.count("input"), x);
                activate_layer(ls, false);
            }
            continue;
        }
        for (Layer *l : seq_order_[si]) {
            LayerState *ls = st_->layer_state(l);
            if (is_input(ls)) {
                const auto ik = in_keys_.find(ls);
                init_layer(ls, ik == in_keys_.end() ||
                               !ik->second.count("input"), x);
            } else if (!init_fold_.count(ls)) {
                init_layer(ls, true, x);
            }
            activate_layer(ls, false);
            update_neurons({ls}, main_of(ls), x);
        }
    }
}

// 4. neuron updates of the parallel structures: the Izhikevich layers that
// share a stream in one launch.
void Runner::neurons(const StepCtx &x) {
    std::map<int, std::vector<LayerState *>> by_stream;
    std::vector<int> order;
    for (auto &lsp : st_->layers()) {
        LayerState *ls = lsp.get();
        if (in_seq_.count(ls->layer)) continue;
        const int m = main_of(ls);
        if (!by_stream.count(m)) order.push_back(m);
        by_stream[m].push_back(ls);
    }
    for (int m : order) update_neurons(by_stream[m], m, x);
}

// 5. weight updates; a plastic conductance connection's is deferred into
// its next activation.
void Runner::updates() {
    if (!learning_) return;
    for (auto &csp : st_->conns()) {
        ConnState &cs = *csp;
        if (cs.update_kind < 0) continue;
        if (fusable(cs)) {
            cs.pending_update = true;
            continue;
        }
        LayerState *to = cs.to;
        hipStream_t s = sched().begin(
            main_of(to), {res(RES_SYN, &cs), res(RES_TSS, &cs),
                          res(RES_OUT, to), res(RES_STATE, to)},
            {res(RES_SYN, &cs)});
        kt_.time("iz_update", [&] { launch_update(*st_, cs, s); });
        sched().end();
    }
}

// 6. output: copies into pinned host buffers on the layers' streams; the
// callbacks run once the copies have landed (output_deliver).
bool Runner::output_issue(long it) {
    if (!(it % env_rate_ == 0 && !suppress_ && !out_keys_.empty()))
        return false;
    for (auto &kv : out_keys_) {
        LayerState *ls = kv.first;
        for (auto &k : kv.second) {
            HostBuf &b = *out_bufs_.at({ls, k});
            const void *src = k == "output" ? ls->output.dev
                            : k == "input" ? ls->input.dev
                            : ls->var(k).dev;
            const Res r = k == "output" ? res(RES_OUT, ls)
                        : k == "input" ? res(RES_IN, ls)
                        : res(RES_STATE, ls);
            hipStream_t s = sched().begin(main_of(ls), {r}, {});
            check_hip(sg_hip_copy_u32(src, b.dp, ls->layer->size, s),
                      "output copy");
            sched().end();
            b.record(s);
        }
    }
    return true;
}

void Runner::output_deliver() {
    for (auto &kv : out_bufs_) kv.second->wait();
    for (auto &m : modules_) {
        if (m.start_delay > 0) continue;
        if (!(m.cutoff == 0 || m.curr < m.cutoff)) continue;
        if (m.curr % m.rate != 0) continue;
        for (auto &ml : m.layers) {
            if (!(ml.io & IO_OUTPUT)) continue;
            for (auto &k : ml.out_keys)
                ml.fn(ml.id, ml.ls->layer->size,
                      out_bufs_.at({ml.ls, k})->p);
        }
    }
}

// modules cycle (io/module.cpp:98-108)
void Runner::cycle_modules() {
    for (auto &m : modules_) {
        if (m.start_delay <= 0) {
            if (m.cutoff == 0 || m.curr < m.cutoff) ++m.curr;
            else if (m.cutoff != 0 && m.curr == m.cutoff) ++m.curr;
        } else {
            --m.start_delay;
        }
    }
}

// ---- captured steps -------------------------------------------------------------
// Everything a captured run of steps bakes in: the device buffers (their
// epoch), the streams, and the host-side state that changes from step to
// step and returns after an even number of steps (each connection's
// time-since-spike buffer and deferred update, the game-of-life word
// buffers).
std::vector<uint64_t> Runner::graph_key(int block) const {
    std::vector<uint64_t> k{st_->device_epoch(), (uint64_t)block,
                            (uint64_t)learning_, (uint64_t)nstreams_};
    for (auto &kv : main_) k.push_back((uint64_t)kv.second);
    for (auto &kv : side_) k.push_back((uint64_t)kv.second);
    for (auto &csp : st_->conns())
        k.push_back((uint64_t)csp->tss_cur |
                    ((uint64_t)csp->pending_update << 1) |
                    ((uint64_t)csp->fused_gol << 2));
    for (auto &kv : gol_fast_)
        k.push_back((uint64_t)(uintptr_t)kv.first->output.dev);
    return k;
}

// The graph of `block` steps from the current host state: captured once
// (steps issued as in an eager run, every draw's step read from the device
// counter, which the graph advances by `block` at its end) and kept with
// the state.
StepGraph *Runner::cached_graph(int block) const {
    const std::vector<uint64_t> key = graph_key(block);
    for (auto &g : C_->graphs)
        if (g->key == key) return g.get();
    return nullptr;
}

StepGraph *Runner::step_graph(int block) {
    if (StepGraph *g = cached_graph(block)) return g;
    const std::vector<uint64_t> key = graph_key(block);
    Sched &S = sched();
    S.barrier();
    hipStream_t s0 = S.stream(0);
    std::unique_ptr<StepGraph> g(new StepGraph());
    check_hip(hipStreamBeginCapture(s0, hipStreamCaptureModeThreadLocal),
              "hipStreamBeginCapture");
    try {
        S.fork();
        S.reset();
        for (int j = 0; j < block; ++j) {
            const StepCtx x{(uint64_t)j, C_->d_step, false};
            pre_input(x);
            post_input(x);
            neurons(x);
            updates();
        }
        S.join();
        hipStream_t s = S.begin(0, {}, {res(RES_CTR, C_->d_step)});
        check_hip(sg_hip_add_u64(C_->d_step, (uint64_t)block, s),
                  "step counter");
        S.end();
        check_hip(hipStreamEndCapture(s0, &g->graph), "hipStreamEndCapture");
    } catch (...) {
        hipGraph_t partial = nullptr;
        (void)hipStreamEndCapture(s0, &partial);
        if (partial) (void)hipGraphDestroy(partial);
        S.reset();
        throw;
    }
    S.reset();
    check_hip(hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0),
              "hipGraphInstantiate");
    g->key = key;
    if (C_->graphs.size() >= 4) C_->graphs.erase(C_->graphs.begin());
    C_->graphs.push_back(std::move(g));
    return C_->graphs.back().get();
}

Props *Runner::run() {
    // Devices: GPUs are 0..n-1, the host is n (resource_manager.h:53).
    int ngpu = sg_hip_device_count();
    if (ngpu <= 0)
        throw Error("No HIP device available: this library runs the "
                    "simulation on an MI355X GPU");
    std::vector<std::string> devs;
    if (args_.has("devices")) devs.push_back(args_.get("devices"));
    for (auto &d : args_.get_array("devices")) devs.push_back(d);
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
    seq_order_.resize(net_->structures().size());
    for (size_t i = 0; i < net_->structures().size(); ++i) {
        const Structure &st = *net_->structures()[i];
        if (st.cluster == CLUSTER_PARALLEL) continue;
        if (st.cluster == CLUSTER_FEEDFORWARD) check_feedforward(st);
        seq_order_[i] = sequential_order(st);
        for (auto &l : st.layers) in_seq_.insert(l.get());
    }

    // "compact records": "false" keeps every per-synapse field in the
    // Izhikevich records uploaded by this run.
    st_->to_device(args_.get_bool("compact records", true));
    C_ = &cache_of(*st_);
    if (!C_->d_step) {
        check_hip(hipMalloc((void **)&C_->d_step, 8), "hipMalloc");
        check_hip(hipHostMalloc((void **)&C_->h_step, 8,
                                hipHostMallocDefault), "hipHostMalloc");
    }

    learning_ = args_.get_bool("learning flag", true);
    suppress_ = args_.get_bool("suppress output", false);
    env_rate_ = std::max(1, args_.get_int("environment rate", 1));
    iterations_ = args_.has("iterations") ? args_.get_int("iterations", 1) : 0;
    kt_.on = args_.get_bool("kernel timing", false);
    step0_ = st_->steps_done;

    setup_modules();
    setup_plans();
    struct GolReset {
        std::map<LayerState *, ConnState *> &m;
        ~GolReset() { for (auto &kv : m) kv.second->fused_gol = false; }
    } gol_reset{gol_fast_};
    setup_streams();
    Sched &S = sched();
    hipStream_t s0 = S.stream(0);

    // Run start: the engine's streams follow everything queued before on
    // the state (uploads ran synchronously); planes of the Izhikevich
    // layers' newest bits; internal dendritic registers at their nodes'
    // init values; the last fed input cleared.
    S.barrier();
    for (auto &lsp : st_->layers()) {
        LayerState *ls = lsp.get();
        if (ls->d_plane) {
            const Res r = res(RES_OUT, ls);
            hipStream_t s = S.begin(0, {r}, {r});
            check_hip(sg_hip_pack_plane(ls->output.du(), ls->layer->size, 0,
                                        ls->d_plane, s), "pack plane");
            S.end();
        }
        const int n = ls->layer->size;
        for (size_t i = 1; i < ls->layer->dendrites.size(); ++i) {
            const Dendrite &d = *ls->layer->dendrites[i];
            const Res r = res(RES_IN, ls);
            hipStream_t s = S.begin(main_of(ls), {r}, {r});
            check_hip(sg_hip_set_f32(ls->input.df() + (size_t)d.reg * n,
                                     d.init, n, 1, s), "dendrite init");
            S.end();
        }
    }

    // With "step graphs": "true", runs without environment I/O capture their
    // steady-state steps (from the second on, all but the last) into HIP
    // graphs of `block` steps -- when at most two streams take part (one
    // structure and its activator stream): a capture over three streams
    // (two structures) ended the process with a segmentation fault on the
    // GPU host once, cause not found, so such runs stay eager.  Off by
    // default: on the benchmark networks the GPU, not the host's launches,
    // bounds the step (the same step time with and without graphs), and a
    // capture costs milliseconds.
    const bool graphs = modules_.empty() && !kt_.on && !any_second_order_ &&
                        nstreams_ <= 2 &&
                        args_.get_bool("step graphs", false) &&
                        iterations_ > 0;
    int block = std::max(2, args_.get_int("graph steps", 8));
    block += block & 1;
    // captured steps launch one game-of-life step each
    if (graphs) gol_multi_.clear();

    const auto t0 = std::chrono::steady_clock::now();
    long it = 0, done = 0, graph_launches = 0;
    bool interrupted = false, pre_issued = false, counter_set = false;
    bool aligned = false;
    auto ctx = [&](long i) {
        return StepCtx{global(i), nullptr,
                       iterations_ != 0 && i + 1 == iterations_};
    };
    while (iterations_ == 0 || it < iterations_) {
        // A graph kept from an earlier run may have been captured one step
        // off (the host state alternates from step to step): one more
        // eager step aligns the run with it instead of capturing anew.
        const bool graph_now = graphs && it >= 1 && iterations_ - it > block;
        if (graph_now && !aligned && !C_->graphs.empty() &&
            !cached_graph(block) && iterations_ - it > block + 1) {
            aligned = true;
        } else if (graph_now) {
            StepGraph *g = step_graph(block);
            S.barrier();
            if (!counter_set) {
                *C_->h_step = global(it);
                check_hip(hipMemcpyAsync(C_->d_step, C_->h_step, 8,
                                         hipMemcpyHostToDevice, s0),
                          "step counter");
                counter_set = true;
            }
            while (iterations_ - it > block) {
                check_hip(hipGraphLaunch(g->exec, s0), "hipGraphLaunch");
                ++graph_launches;
                it += block;
                done += block;
                if (g_interrupt) break;
            }
            S.barrier();
            check_hip(hipGetLastError(), "step");
            if (g_interrupt) {
                interrupted = true;
                break;
            }
            continue;
        }
        const StepCtx x = ctx(it);
        gol_rem_ = iterations_ != 0 ? iterations_ - it : (1L << 40);
        if (!pre_issued) pre_input(x);
        pre_issued = false;
        input_phase();
        post_input(x);
        neurons(x);
        updates();
        const bool out = output_issue(it);
        ++done;
        // With environment I/O, the next step's pre-input launches go out
        // before the host runs this step's output callbacks (and the next
        // input callbacks), so the GPU works while the host does.
        // An interrupt ends the run at a step that no game-of-life launch
        // has run past.
        const bool more = (iterations_ == 0 || it + 1 < iterations_) &&
                          (!g_interrupt || gol_pending());
        if (more && !modules_.empty()) {
            pre_input(ctx(it + 1));
            pre_issued = true;
        }
        if (out) output_deliver();
        cycle_modules();
        check_hip(hipGetLastError(), "step");
        ++it;
        if (g_interrupt && !pre_issued && !gol_pending()) {
            interrupted = true;
            break;
        }
    }
    // Leave double-buffered game-of-life words in the state's own buffer.
    S.barrier();
    for (auto &kv : gol_fast_) {
        LayerState *ls = kv.first;
        if (ls->output.dev != gol_home_[ls]) {
            check_hip(hipMemcpyAsync(gol_home_[ls], ls->output.dev,
                                     ls->output.bytes(),
                                     hipMemcpyDeviceToDevice, s0),
                      "gol restore");
            C_->gol_spare[ls] = ls->output.du();
            ls->output.dev = gol_home_[ls];
        }
    }
    S.join();
    check_hip(hipStreamSynchronize(s0), "final sync");
    S.reset();
    st_->steps_done += (uint64_t)done;
    Props timing;
    kt_.report(timing);
    const double secs = std::chrono::duration<double>(
        std::chrono::steady_clock::now() - t0).count();

    // Report (report.cpp:8-16, engine.cpp:584-590).
    Props *r = new Props();
    r->set("iterations", std::to_string(done));
    if (!gol_fast_.empty())
        r->set("game of life launches", std::to_string(gol_launches_));
    r->set("total time", std::to_string((float)secs));
    r->set("average time", std::to_string((float)(secs / (double)done)));
    r->set("fps", std::to_string((float)((double)done / secs)));
    r->set("network bytes", std::to_string(st_->network_bytes()));
    size_t buf_bytes = 0;
    for (auto &kv : in_bufs_) buf_bytes += (size_t)kv.second->b[0].n * 4;
    for (auto &kv : out_bufs_) buf_bytes += (size_t)kv.second->n * 4;
    r->set("state buffer bytes",
           std::to_string(any_input_key() ? buf_bytes : 0));
    r->set("engine buffer bytes", std::to_string(buf_bytes));
    r->set_child("args", args_);
    r->set("interrupted", interrupted ? "true" : "false");
    // 2: the input-writing activators ran on a second stream beside the
    // conductance passes; 1: every activator on its structure's stream.
    r->set("activator streams", any_split_ ? "2" : "1");
    // streams used in all (structures, activator streams, IO)
    r->set("streams", std::to_string(nstreams_));
    // captured graphs launched, and the steps they ran
    r->set("step graphs", std::to_string(graph_launches));
    r->set("graph steps", std::to_string(graph_launches * block));
    // Izhikevich connections whose records are compact (sg_hip.h SG_IZC_*).
    int compact = 0;
    for (auto &csp : st_->conns()) compact += csp->compact ? 1 : 0;
    r->set("compact records", std::to_string(compact));
    Props mem;
    mem.set("device", std::to_string(device));
    mem.set("bytes", std::to_string(st_->network_bytes()));
    r->add_to_child_array("memory usage", mem);
    if (kt_.on) r->set_child("kernel timing", timing);
    return r;
}

}  // namespace

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
    Runner runner(net_, env_, state_, args);
    return runner.run();
}

}  // namespace sg
