// Simulation state: every neuron and synapse array of a network, kept as a
// host mirror in the reference's host layout (what the getters and save
// files expose) and, once the engine runs, as a device twin resident in HBM
// in the layout the kernels read.
//
// Data lives on exactly one side at a time, with the reference's transfer
// semantics (state/state.cpp:259-323): the engine moves everything to the
// device at the start of a run; any getter moves everything back to the host
// and hands out a writable view that stays valid until the next run.
#ifndef SG_STATE_H
#define SG_STATE_H

#include <cstdint>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "network.h"
#include "props.h"
#include "sg_hip.h"

namespace sg {

enum class Elem { F32, I32, U32 };
enum class Model { IZHIKEVICH, GAME_OF_LIFE, RATE_ENCODING, RELAY };

// One named array.  `host` is the reference's host layout.  On the device
// the array is either a copy or, for per-destination synapse arrays
// ([rows = destination slots][cols = K] on the host), the 64-destination
// tiled layout the synapse kernels stream.
struct Array {
    std::string key;
    Elem elem = Elem::F32;
    std::vector<uint32_t> host;
    void *dev = nullptr;
    size_t dev_elems = 0;   // elements in the device twin (may be < host)
    int rows = 0, cols = 0;
    bool tiled = false;

    float *f() { return reinterpret_cast<float *>(host.data()); }
    int32_t *i() { return reinterpret_cast<int32_t *>(host.data()); }
    uint32_t *u() { return host.data(); }
    size_t size() const { return host.size(); }
    size_t bytes() const { return host.size() * 4; }
    float *df() const { return static_cast<float *>(dev); }
    int32_t *di() const { return static_cast<int32_t *>(dev); }
    uint32_t *du() const { return static_cast<uint32_t *>(dev); }
};

struct LayerState {
    Layer *layer = nullptr;
    Model model = Model::RELAY;
    int out_type = SG_OUT_FLOAT;
    int history = 1;
    Array input, output;
    std::map<std::string, Array> vars;   // neuron variables by key
    // model parameters
    int survival_min = 2, survival_max = 3, birth_min = 3, birth_max = 3;
    bool ramp = false;

    // Izhikevich layers: plane of the newest spike bits (one bit per
    // neuron), written by the neuron kernel for delay-0 base activators.
    uint64_t *d_plane = nullptr;

    Array &var(const std::string &k) { return vars.at(k); }
};

struct ConnState {
    Connection *conn = nullptr;
    LayerState *from = nullptr, *to = nullptr;
    std::map<std::string, Array> arrays;   // "weights", "delays", ...
    bool iz = false;          // Izhikevich weight matrix
    bool direct = false;      // base activator even on an Izhikevich layer
    int iz_kind = -1;         // SG_IZ_* conductance activator
    int update_kind = -1;     // 0 = update_iz_add, 1 = update_iz_sub
    float baseline_conductance = 1.0f, learning_rate = 0.1f;
    bool stp_flag = true;
    int num_dest = 0, k_per_dest = 0;
    // device-only derived arrays
    float *d_record = nullptr;        // interleaved per-synapse record
    bool packed = false;              // iz arrays live in d_record
    int record_fields = 0;            // fields d_record was allocated with
    // Compact record (sg_hip.h SG_IZC_*): presynaptic trace and short-term
    // plasticity kept per source in d_src_stp ({R, w, pre, -} per source).
    bool compact = false;
    float *d_src_stp = nullptr;
    uint8_t *d_delays = nullptr;      // delay & 31, synapse layout
    uint8_t *d_rev_delays = nullptr;  // [source][32] last visits, reversed
    uint8_t *d_visits = nullptr;      // min(32, visits per source)
    int32_t *d_tss[2] = {nullptr, nullptr};
    uint64_t *d_plane = nullptr;      // spike plane of a BIT source (base)
    int32_t *d_sparse_from = nullptr; // SG_SPARSE source indices, tiled
    int32_t *d_tile_counts = nullptr; // SG_SPARSE {fewest, most} per tile
    int tss_cur = 0;
    bool pending_update = false;
    bool fused_gol = false;   // handled by sg_hip_gol_step
    bool finite_weights = false;  // every weight finite at upload

    Array &arr(const std::string &k) { return arrays.at(k); }
    bool has(const std::string &k) const { return arrays.count(k) > 0; }
};

// Engine resources kept from one run to the next (streams, events, spare
// buffers, captured step graphs): owned by the state they were made for,
// built and used by the engine.
struct EngineCache {
    virtual ~EngineCache() {}
};

// Warnings of state construction (the reference's LOG_WARNING), shown
// unless set_warnings(false) or set_suppress_output(true) turned them off.
void set_host_messages(bool warnings, bool suppress);
void warn(const std::string &msg);

class State {
  public:
    explicit State(Network *net);
    ~State();

    Network *network() const { return net_; }
    // Host views (moves data to the host first).
    Array *neuron_data(Layer *layer, const std::string &key);
    Array *weight_matrix(Connection *conn, const std::string &key);

    // Applies deferred updates (see pending_update), then moves every array
    // to the host.
    void to_host();
    void download_compact(ConnState &cs);
    // allow_compact: tiled Izhikevich conductance connections whose
    // per-source values agree get the compact record.
    void to_device(bool allow_compact = true);
    bool on_host() const { return on_host_; }

    std::vector<std::unique_ptr<LayerState>> &layers() { return layers_; }
    std::vector<std::unique_ptr<ConnState>> &conns() { return conns_; }
    LayerState *layer_state(Layer *l) const;
    ConnState *conn_state(Connection *c) const;

    size_t network_bytes() const;
    // checkpoint (state/state.cpp:365-447)
    void save(const std::string &path);
    void load(const std::string &path);

    // device scratch for layout changes
    void *scratch(size_t bytes);
    // iSTDP tables for update_iz_sub, computed once per state on the host
    const float *d_src_table() const { return d_tables_; }
    const float *d_dst_table() const { return d_tables_ + 32; }
    void ensure_device();
    // Incremented whenever device buffers are (re)filled from the host:
    // what the engine captured before is stale then.
    uint64_t device_epoch() const { return device_epoch_; }

    // Steps simulated on this state so far, over all runs: the step index
    // that keys random init values.
    uint64_t steps_done = 0;
    // Engine resources of this state (see EngineCache).
    std::unique_ptr<EngineCache> engine_cache;

  private:
    void build_layer(Layer *l);
    void build_connection(Connection *c);
    void init_weights(ConnState &cs);
    void process_iz_matrix(ConnState &cs);
    void compute_indices(ConnState &cs);
    void sparsify(ConnState &cs);
    void randomize_projection(ConnState &cs);
    void adjust_sparse_indices(ConnState &cs);
    void compute_distances(ConnState &cs, float from_spacing,
                           float to_spacing, float x_offset, float y_offset);
    void upload_conn_derived(ConnState &cs);
    void upload(Array &a);
    void download(Array &a);

    Network *net_;
    std::vector<std::unique_ptr<LayerState>> layers_;
    std::vector<std::unique_ptr<ConnState>> conns_;
    bool on_host_ = true;
    bool device_ready_ = false;
    uint64_t device_epoch_ = 0;
    void *scratch_ = nullptr;
    size_t scratch_bytes_ = 0;
    float *d_tables_ = nullptr;
};

// iSTDP factors of update_iz_sub for time-since-spike values 0..31, with the
// reference's expressions (izhikevich_attributes.cpp:454-472).
void istdp_tables(float src[32], float dst[32]);

// Output type of a neural model name; throws for unsupported models.
Model model_from(const std::string &name);
int model_out_type(Model m);

void check_hip(int code, const char *what);

}  // namespace sg

#endif
