#include "state.h"
#include "engine.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <functional>
#include <random>
#include <sstream>

#include "callbacks.h"

namespace sg {

static bool g_host_warnings = true, g_host_suppress = false;

void set_host_messages(bool warnings, bool suppress) {
    g_host_warnings = warnings;
    g_host_suppress = suppress;
}

void warn(const std::string &msg) {
    if (g_host_warnings && !g_host_suppress)
        fprintf(stderr, "%s\n", msg.c_str());
}

void check_hip(int code, const char *what) {
    if (code != 0)
        throw Error(std::string(what) + ": " + hipGetErrorString((hipError_t)code));
}

Model model_from(const std::string &name) {
    if (name == "izhikevich") return Model::IZHIKEVICH;
    if (name == "game_of_life") return Model::GAME_OF_LIFE;
    if (name == "rate_encoding") return Model::RATE_ENCODING;
    if (name == "relay") return Model::RELAY;
    throw Error("Unrecognized neural model \"" + name + "\"");
}

int model_out_type(Model m) {
    return (m == Model::IZHIKEVICH || m == Model::GAME_OF_LIFE) ? SG_OUT_BIT
                                                                : SG_OUT_FLOAT;
}

// Weight-initialisation generator.  The reference draws from an unseeded
// std::mt19937 (util/tools.h:8); this one is seeded so that random weight
// configs are reproducible (SYNGEN_SEED overrides the seed).
static std::mt19937 &rng() {
    static std::mt19937 gen([] {
        const char *s = std::getenv("SYNGEN_SEED");
        return s ? (unsigned)std::strtoul(s, nullptr, 10) : 5489u;
    }());
    return gen;
}

static Array make_array(const std::string &key, Elem e, size_t n,
                        uint32_t fill = 0) {
    Array a;
    a.key = key;
    a.elem = e;
    a.host.assign(n, fill);
    return a;
}

static uint32_t fbits(float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    return u;
}

// ---------------------------------------------------------------------------
// Izhikevich parameter table (izhikevich_attributes.cpp:28-35).
struct IzParams { float a, b, c, d; };

static IzParams iz_params(const std::string &s) {
    if (s == "default") return {0.02f, 0.2f, -70.0f, 2.0f};
    if (s == "regular") return {0.02f, 0.2f, -65.0f, 8.0f};
    if (s == "bursting") return {0.02f, 0.2f, -55.0f, 4.0f};
    if (s == "chattering") return {0.02f, 0.2f, -50.0f, 2.0f};
    if (s == "fast") return {0.1f, 0.2f, -65.0f, 2.0f};
    if (s == "low_threshold") return {0.02f, 0.25f, -65.0f, 2.0f};
    if (s == "thalamo_cortical") return {0.02f, 0.25f, -65.0f, 0.05f};
    if (s == "resonator") return {0.1f, 0.26f, -65.0f, 2.0f};
    throw Error("Unrecognized parameter string: " + s);
}

State::State(Network *net) : net_(net) {
    for (auto *l : net->layers()) build_layer(l);
    for (auto *l : net->layers())
        for (auto *c : l->inputs) build_connection(c);
    // Sparse wrapped indices are remapped only now, after every delay was
    // initialised from the pre-wrap ones (state.cpp:108-115).
    for (auto &cs : conns_)
        if (cs->conn->type == SG_SPARSE) adjust_sparse_indices(*cs);
    // Index arrays needed only while weights and delays are initialised are
    // dropped once every connection is built, so they are not part of a
    // state file (state.cpp:108-119, weight_matrix.cpp:345-353).
    for (auto &cs : conns_)
        for (const char *k : {"used", "to row indices", "to column indices"})
            cs->arrays.erase(k);
}

State::~State() {
    // engine streams and graphs go first: they may reference the buffers
    engine_cache.reset();
    auto free_dev = [](void *p) { if (p) (void)hipFree(p); };
    for (auto &ls : layers_) {
        free_dev(ls->input.dev);
        free_dev(ls->output.dev);
        free_dev(ls->d_plane);
        for (auto &kv : ls->vars) free_dev(kv.second.dev);
    }
    for (auto &cs : conns_) {
        for (auto &kv : cs->arrays) free_dev(kv.second.dev);
        free_dev(cs->d_record);
        free_dev(cs->d_src_stp);
        free_dev(cs->d_delays);
        free_dev(cs->d_rev_delays);
        free_dev(cs->d_visits);
        free_dev(cs->d_tss[0]);
        free_dev(cs->d_tss[1]);
        free_dev(cs->d_plane);
        free_dev(cs->d_sparse_from);
        free_dev(cs->d_tile_counts);
    }
    free_dev(scratch_);
    free_dev(d_tables_);
}

void State::build_layer(Layer *l) {
    std::unique_ptr<LayerState> ls(new LayerState());
    ls->layer = l;
    ls->model = model_from(l->neural_model);
    ls->out_type = model_out_type(ls->model);
    const int per_word = ls->out_type == SG_OUT_BIT ? 32 : 1;
    ls->history = 1 + l->max_out_delay() / per_word;
    const int n = l->size;
    // one input register per dendritic node with children of its own
    // (attributes.cpp:23-28); register 0 is the neurons' input
    ls->input = make_array("input", Elem::F32, (size_t)l->registers * n);
    ls->output = make_array("output", Elem::U32, (size_t)ls->history * n);

    if (ls->model == Model::IZHIKEVICH) {
        for (const char *k : {"ampa", "nmda", "gabaa", "gabab", "mult",
                              "dopamine", "voltage", "recovery", "post trace",
                              "a", "b", "c", "d"})
            ls->vars[k] = make_array(k, Elem::F32, n);
        ls->vars["time since spike"] =
            make_array("time since spike", Elem::I32, n, 32);
        float *as = ls->var("a").f(), *bs = ls->var("b").f();
        float *cs = ls->var("c").f(), *ds = ls->var("d").f();
        const std::string p = l->param("params", "regular");
        std::uniform_real_distribution<float> u01(0.0f, 1.0f);
        for (int i = 0; i < n; ++i) {
            if (p == "random positive") {
                as[i] = 0.02f;
                bs[i] = 0.2f;
                cs[i] = (float)(-65.0 + 15.0 * std::pow(u01(rng()), 2));
                ds[i] = (float)(8.0 - 6.0 * std::pow(u01(rng()), 2));
            } else if (p == "random negative") {
                as[i] = (float)(0.02 + 0.08 * u01(rng()));
                bs[i] = (float)(0.25 - 0.05 * u01(rng()));
                cs[i] = -65.0f;
                ds[i] = 2.0f;
            } else {
                IzParams q = iz_params(p);
                as[i] = q.a; bs[i] = q.b; cs[i] = q.c; ds[i] = q.d;
            }
        }
        // Settle to a resting point (izhikevich_attributes.cpp:628-649).
        float *v = ls->var("voltage").f(), *u = ls->var("recovery").f();
        for (int j = 0; j < n; ++j) {
            float vv = cs[j], r = bs[j] * cs[j], dv, dr;
            const float a = as[j], b = bs[j];
            do {
                dv = (float)((0.04 * (double)vv) * (double)vv +
                             (double)(5.0f * vv) + 140.0 - (double)r);
                vv = vv + dv;
                dr = a * ((b * vv) - r);
                r = r + dr;
            } while ((double)std::fabs(dv) > 0.001 &&
                     (double)std::fabs(dr) > 0.001);
            v[j] = vv;
            u[j] = r;
        }
    } else if (ls->model == Model::GAME_OF_LIFE) {
        ls->survival_min = std::stoi(l->param("survival_min", "2"));
        ls->survival_max = std::stoi(l->param("survival_max", "3"));
        ls->birth_min = std::stoi(l->param("birth_min", "3"));
        ls->birth_max = std::stoi(l->param("birth_max", "3"));
    } else if (ls->model == Model::RELAY) {
        ls->ramp = l->param("ramp", "false") == "true";
    }
    layers_.push_back(std::move(ls));
}

LayerState *State::layer_state(Layer *l) const {
    for (auto &ls : layers_)
        if (ls->layer == l) return ls.get();
    throw Error("Unknown layer " + l->name);
}

ConnState *State::conn_state(Connection *c) const {
    for (auto &cs : conns_)
        if (cs->conn == c) return cs.get();
    throw Error("Unknown connection " + c->name);
}

// ---------------------------------------------------------------------------
// Connections

void State::build_connection(Connection *c) {
    if (c->sparse) {
        c->type = c->dense_type;
        c->num_weights = c->dense_num_weights;
    }
    std::unique_ptr<ConnState> cs(new ConnState());
    cs->conn = c;
    cs->from = layer_state(c->from);
    cs->to = layer_state(c->to);
    const int nw = c->num_weights;
    switch (c->type) {
        case SG_FULLY_CONNECTED:
            cs->num_dest = c->to->size;
            cs->k_per_dest = c->from->size;
            break;
        case SG_SUBSET:
            cs->num_dest = c->subset.to_size();
            cs->k_per_dest = c->subset.from_size();
            break;
        case SG_ONE_TO_ONE:
            cs->num_dest = c->from->size;
            cs->k_per_dest = 1;
            break;
        default:
            cs->num_dest = c->to->size;
            cs->k_per_dest = c->arbor.total();
    }
    cs->arrays["weights"] = make_array("weights", Elem::F32, nw);
    conns_.push_back(std::move(cs));
    ConnState &s = *conns_.back();

    // Activator / updater selection (izhikevich_attributes.cpp:368-398,
    // 522-539; kernel.cpp:222-243 for every other model).
    s.iz = s.to->model == Model::IZHIKEVICH;
    if (s.iz) {
        s.direct = c->param("direct", "false") == "true";
        if (!s.direct) {
            switch (c->opcode) {
                case SG_ADD: s.iz_kind = SG_IZ_ADD; break;
                case SG_SUB: s.iz_kind = SG_IZ_SUB; break;
                case SG_MULT: s.iz_kind = SG_IZ_MULT; break;
                case SG_REWARD: s.iz_kind = SG_IZ_REWARD; break;
                case SG_GAP: s.iz_kind = SG_IZ_GAP; break;
                default: throw Error("Unimplemented connection type!");
            }
        }
        if (c->plastic) {
            if (c->convolutional) throw Error("Unimplemented connection type!");
            if (c->opcode == SG_ADD) s.update_kind = 0;
            else if (c->opcode == SG_SUB) s.update_kind = 1;
            else throw Error("Unimplemented connection type!");
        }
    }
    if (s.iz_kind == SG_IZ_GAP && c->from->size > c->to->size)
        throw Error("Gap junctions read the destination layer's voltages at "
                    "source indices; the source layer must not be larger");
    if (!s.iz || s.direct) {
        if (c->opcode == SG_REWARD || c->opcode == SG_MODULATE ||
            c->opcode == SG_GAP)
            throw Error("Opcode " + c->param("opcode", "") +
                        " has no base aggregator");
    }

    if (c->randomized) randomize_projection(s);
    init_weights(s);
    if (c->sparse) sparsify(s);
    if ((c->so_host || c->so_slave) && s.iz && !s.direct)
        throw Error("Second-order connections into an Izhikevich "
                    "conductance path are not supported (" + c->str() + ")");
    // the host's weight copy for its slaves (weight_matrix.cpp:78-80)
    if (c->so_host)
        s.arrays["second order weights"] =
            make_array("second order weights", Elem::F32, c->num_weights);
    if (s.iz) {
        const int n = c->num_weights;
        for (const char *k : {"short trace", "long trace", "presyn trace",
                              "short term R", "short term w",
                              "weight derivatives"})
            s.arrays[k] = make_array(k, Elem::F32, n);
        s.arrays["time since spike"] =
            make_array("time since spike", Elem::I32, n);
        process_iz_matrix(s);
    }
}

// WeightMatrix::sparsify (weight_matrix.cpp:510-663): zero weights become
// unused; each destination row keeps its used entries in order, packed to
// the front of a row of max_nonzero entries (weights padded with 0, index
// arrays with -1, distances and delays with 0), nonzero_counts per row.
// The connection turns SG_SPARSE (connection.cpp:219-222).
void State::sparsify(ConnState &cs) {
    Connection &c = *cs.conn;
    compute_indices(cs);
    const int rows = c.to->size;
    const int cols = c.num_weights / rows;
    const float *w = cs.arr("weights").f();
    int32_t *used = cs.arr("used").i();
    int maxnz = 0;
    for (int r = 0; r < rows; ++r) {
        int nz = 0;
        for (int k = 0; k < cols; ++k) {
            const long i = (long)r * cols + k;
            if (w[i] == 0.0f) used[i] = 0;
            nz += used[i];
        }
        maxnz = std::max(maxnz, nz);
    }
    if (maxnz == 0) {
        warn("Warning in weight config for " + c.str() + ":\n" +
             "    Attempted to sparsify empty matrix!");
        maxnz = 1;
        for (int r = 0; r < rows; ++r) used[(long)r * cols] = 1;
    }
    const long snw = (long)maxnz * rows;
    Array counts = make_array("nonzero counts", Elem::I32, rows, 0);
    int32_t *cnt = counts.i();
    auto compact = [&](const char *key, uint32_t pad) {
        if (!cs.has(key)) return;
        Array &a = cs.arr(key);
        Array out = make_array(key, a.elem, snw, pad);
        for (int r = 0; r < rows; ++r) {
            int n = 0;
            for (int k = 0; k < cols; ++k) {
                const long i = (long)r * cols + k;
                if (used[i]) out.host[(long)r * maxnz + n++] = a.host[i];
            }
        }
        a = std::move(out);
    };
    for (int r = 0; r < rows; ++r) {
        int n = 0;
        for (int k = 0; k < cols; ++k) n += used[(long)r * cols + k];
        cnt[r] = n;
    }
    compact("weights", 0u);
    for (const char *k : {"from row indices", "from column indices",
                          "from indices", "to row indices",
                          "to column indices", "to indices"})
        compact(k, (uint32_t)-1);
    compact("distances", 0u);
    compact("delays", 0u);
    compact("used", 0u);
    cs.arrays["nonzero counts"] = std::move(counts);
    c.num_weights = (int)snw;
    c.type = SG_SPARSE;
    cs.num_dest = rows;
    cs.k_per_dest = maxnz;
}

// WeightMatrix::adjust_sparse_indices (weight_matrix.cpp:253-343): wrapped
// arborized sources are remapped into the layer; without wrap, sources out
// of bounds become unused and each row's index entries shift left over
// them (the weights stay where they are, as in the reference).
void State::adjust_sparse_indices(ConnState &cs) {
    Connection &c = *cs.conn;
    const int rows = c.to->size, maxnz = cs.k_per_dest;
    const int fr_n = c.from->rows, fc_n = c.from->cols;
    const bool wrap = c.arbor.wrap;
    int32_t *fr = cs.arr("from row indices").i();
    int32_t *fc = cs.arr("from column indices").i();
    int32_t *fi = cs.arr("from indices").i();
    int32_t *tr = cs.arr("to row indices").i();
    int32_t *tc = cs.arr("to column indices").i();
    int32_t *ti = cs.arr("to indices").i();
    int32_t *used = cs.arr("used").i();
    int32_t *cnt = cs.arr("nonzero counts").i();
    for (int r = 0; r < rows; ++r)
        for (int k = 0; k < maxnz; ++k) {
            const long i = (long)r * maxnz + k;
            if (!used[i]) continue;
            if (wrap) {
                fr[i] = (fr[i] % fr_n + fr_n) % fr_n;
                fc[i] = (fc[i] % fc_n + fc_n) % fc_n;
                fi[i] = fr[i] * fc_n + fc[i];
            } else if (fr[i] < 0 || fr[i] >= fr_n || fc[i] < 0 ||
                       fc[i] >= fc_n) {
                used[i] = 0;
            }
        }
    if (wrap) return;
    for (int r = 0; r < rows; ++r) {
        const long off = (long)r * maxnz;
        long cur = off, dst = off;
        for (int k = 0; k < cnt[r]; ++k, ++cur) {
            if (!used[cur]) continue;
            if (cur != dst) {
                fr[dst] = fr[cur]; fc[dst] = fc[cur]; fi[dst] = fi[cur];
                tr[dst] = tr[cur]; tc[dst] = tc[cur]; ti[dst] = ti[cur];
                used[dst] = 1;
                used[cur] = 0;
            }
            ++dst;
        }
        cnt[r] = (int)(dst - off);
    }
}

// Host-layout weight index -> (slot, k) and the serial enumeration used for
// indices, distances and the time-since-spike visit ranks.
template <class F>
static void enumerate_synapses(const Connection &c, int num_dest, F &&f) {
    // f(slot, to_index, k, from_row, from_col, pre_wrap_row, pre_wrap_col,
    //   host_weight_index)
    const int fr_n = c.from->rows, fc_n = c.from->cols;
    auto wrapi = [](int x, int n) { return ((x % n) + n) % n; };
    for (int slot = 0; slot < num_dest; ++slot) {
        if (c.type == SG_FULLY_CONNECTED) {
            for (int k = 0; k < c.from->size; ++k)
                f(slot, slot, k, k / fc_n, k % fc_n, k / fc_n, k % fc_n,
                  (long)slot * c.from->size + k);
        } else if (c.type == SG_ONE_TO_ONE) {
            f(slot, slot, 0, slot / fc_n, slot % fc_n, slot / fc_n,
              slot % fc_n, (long)slot);
        } else if (c.type == SG_SUBSET) {
            const SubsetConfig &s = c.subset;
            const int sub_cols = s.to_col_end - s.to_col_start;
            const int to = (slot / sub_cols + s.to_row_start) * c.to->cols +
                           slot % sub_cols + s.to_col_start;
            int k = 0;
            for (int r = s.from_row_start; r < s.from_row_end; ++r)
                for (int cc = s.from_col_start; cc < s.from_col_end; ++cc, ++k)
                    f(slot, to, k, r, cc, r, cc,
                      (long)slot * s.from_size() + k);
        } else {
            const ArborConfig &a = c.arbor;
            const int to_row = slot / c.to->cols, to_col = slot % c.to->cols;
            const long base = c.convolutional ? 0 : (long)slot * a.total();
            int k = 0;
            if (c.type == SG_CONVERGENT) {
                const int sr = to_row * a.row_stride + a.row_spacing * a.row_offset;
                const int sc = to_col * a.col_stride + a.col_spacing * a.col_offset;
                for (int kr = 0; kr < a.row_field; ++kr)
                    for (int kc = 0; kc < a.col_field; ++kc, ++k) {
                        int r = sr + kr * a.row_spacing, cc = sc + kc * a.col_spacing;
                        const int pr = r, pc = cc;
                        if (a.wrap) { r = wrapi(r, fr_n); cc = wrapi(cc, fc_n); }
                        else if (r < 0 || r >= fr_n || cc < 0 || cc >= fc_n)
                            continue;
                        f(slot, slot, k, r, cc, pr, pc, base + k);
                    }
            } else {
                const int r0 = (to_row + a.row_spacing * (-a.row_offset -
                                a.row_field + a.row_stride)) / a.row_stride;
                const int c0 = (to_col + a.col_spacing * (-a.col_offset -
                                a.col_field + a.col_stride)) / a.col_stride;
                const int r1 = r0 + (a.row_spacing * (a.row_field -
                                a.row_stride) / a.row_stride);
                const int c1 = c0 + (a.col_spacing * (a.col_field -
                                a.col_stride) / a.col_stride);
                for (int sr = r0; sr <= r1; sr += a.row_spacing)
                    for (int sc = c0; sc <= c1; sc += a.col_spacing, ++k) {
                        int r = sr, cc = sc;
                        if (a.wrap) { r = wrapi(r, fr_n); cc = wrapi(cc, fc_n); }
                        else if (r < 0 || r >= fr_n || cc < 0 || cc >= fc_n)
                            continue;
                        f(slot, slot, k, r, cc, sr, sc, base + k);
                    }
            }
        }
    }
}

// The same enumeration for a connection's state: SG_SPARSE walks each
// row's compacted list (synapse_kernel.h:311-322), any other type its
// geometry.
template <class F>
static void enumerate_synapses(const ConnState &cs, F &&f) {
    const Connection &c = *cs.conn;
    if (c.type != SG_SPARSE) {
        enumerate_synapses(c, cs.num_dest, f);
        return;
    }
    const int K = cs.k_per_dest;
    auto ints = [&](const char *k) {
        return reinterpret_cast<const int32_t *>(cs.arrays.at(k).host.data());
    };
    const int32_t *cnt = ints("nonzero counts");
    const int32_t *fr = ints("from row indices");
    const int32_t *fc = ints("from column indices");
    for (int slot = 0; slot < cs.num_dest; ++slot)
        for (int k = 0; k < cnt[slot]; ++k) {
            const long wi = (long)slot * K + k;
            f(slot, slot, k, fr[wi], fc[wi], fr[wi], fc[wi], wi);
        }
}

// get_indices (weight_matrix.cpp:1039-1195): per host weight index, the
// source/destination rows, columns and indices (pre-wrap for arborized
// connections) and a used flag; unvisited entries keep -1 / 0.
void State::compute_indices(ConnState &cs) {
    if (cs.has("used")) return;
    Connection &c = *cs.conn;
    const size_t nw = c.num_weights;
    for (const char *k : {"from row indices", "from column indices",
                          "from indices", "to row indices",
                          "to column indices", "to indices"})
        cs.arrays[k] = make_array(k, Elem::I32, nw, (uint32_t)-1);
    cs.arrays["used"] = make_array("used", Elem::I32, nw, 0);
    int32_t *fr = cs.arr("from row indices").i();
    int32_t *fc = cs.arr("from column indices").i();
    int32_t *fi = cs.arr("from indices").i();
    int32_t *tr = cs.arr("to row indices").i();
    int32_t *tc = cs.arr("to column indices").i();
    int32_t *ti = cs.arr("to indices").i();
    int32_t *used = cs.arr("used").i();
    const bool arbor = c.type == SG_CONVERGENT || c.type == SG_DIVERGENT;
    const int ncols = c.from->cols;
    enumerate_synapses(cs, [&](int, int to, int, int r, int col, int pr,
                               int pc, long wi) {
        const int rr = arbor ? pr : r, cc = arbor ? pc : col;
        fr[wi] = rr;
        fc[wi] = cc;
        fi[wi] = rr * ncols + cc;
        tr[wi] = to / c.to->cols;
        tc[wi] = to % c.to->cols;
        ti[wi] = to;
        used[wi] = 1;
    });
}

// WeightMatrix::randomize_projection (weight_matrix.cpp:356-507), before
// the weights are initialised: every destination of a one-to-one
// connection reads one random source neuron; every destination of a
// convergent connection keeps its field's shape, moved so that its first
// visit sits at a random source neuron, and its source indices are row *
// rows + column as the reference writes them (with wrap, the sparse index
// adjustment recomputes them from rows and columns).  The draws come from
// the state's seeded generator, where the reference's is unseeded.
void State::randomize_projection(ConnState &cs) {
    Connection &c = *cs.conn;
    compute_indices(cs);
    int32_t *fr = cs.arr("from row indices").i();
    int32_t *fc = cs.arr("from column indices").i();
    int32_t *fi = cs.arr("from indices").i();
    std::uniform_int_distribution<int> row_d(0, c.from->rows - 1);
    std::uniform_int_distribution<int> col_d(0, c.from->cols - 1);
    if (c.type == SG_ONE_TO_ONE) {
        for (int i = 0; i < c.to->size; ++i) {
            const int r = fr[i] = row_d(rng());
            const int q = fc[i] = col_d(rng());
            fi[i] = r * c.from->cols + q;
        }
        return;
    }
    const int K = cs.k_per_dest;
    for (int i = 0; i < c.to->size; ++i) {
        const int nr = row_d(rng()), nc = col_d(rng());
        const long corner = (long)K * i;
        const int dr = nr - fr[corner], dc = nc - fc[corner];
        for (int k = 0; k < K; ++k) {
            const int r = (fr[corner + k] += dr);
            const int q = (fc[corner + k] += dc);
            fi[corner + k] = r * c.from->rows + q;
        }
    }
}

// get_distances (weight_matrix.cpp:1199-1317).
void State::compute_distances(ConnState &cs, float from_spacing,
                              float to_spacing, float x_offset,
                              float y_offset) {
    if (cs.has("distances")) return;
    compute_indices(cs);
    const size_t nw = cs.conn->num_weights;
    cs.arrays["distances"] = make_array("distances", Elem::F32, nw);
    float *d = cs.arr("distances").f();
    int32_t *fr = cs.arr("from row indices").i();
    int32_t *fc = cs.arr("from column indices").i();
    int32_t *tr = cs.arr("to row indices").i();
    int32_t *tc = cs.arr("to column indices").i();
    for (size_t i = 0; i < nw; ++i) {
        const float f_y = fr[i] * from_spacing, f_x = fc[i] * from_spacing;
        const float t_y = tr[i] * to_spacing + y_offset;
        const float t_x = tc[i] * to_spacing + x_offset;
        const double dx = (double)(t_x - f_x), dy = (double)(t_y - f_y);
        d[i] = (float)std::pow(dx * dx + dy * dy, 0.5);
    }
}

static void fill_fraction(float *w, int n, float fraction,
                          const std::function<float()> &gen) {
    std::uniform_real_distribution<float> u01(0.0f, 1.0f);
    for (int i = 0; i < n; ++i)
        if (fraction == 1.0f || u01(rng()) < fraction) w[i] = gen();
}

// Circular mask (weight_matrix.cpp:749-824).
static void circular_mask(const Props &cfg, float *w, const Connection &c) {
    if (c.type != SG_CONVERGENT)
        throw Error("CircularMaskConfig can only be used on convergent "
                    "arborized connections!");
    const int rf = c.arbor.row_field, cf = c.arbor.col_field;
    const int ks = c.arbor.total();
    float row_radius = cfg.get_float("row radius", 0);
    float col_radius = cfg.get_float("column radius", 0);
    const float row_diameter = cfg.get_float("row diameter", 0);
    const float col_diameter = cfg.get_float("column diameter", 0);
    const float radius = cfg.get_float("radius", 0);
    const float diameter = cfg.get_float("diameter", 0);
    const bool invert = cfg.get_bool("invert", false);
    const float value = cfg.get_float("value", 0.0f);
    if (row_radius <= 0) {
        if (row_diameter > 0) row_radius = row_diameter / 2;
        else if (radius > 0) row_radius = radius;
        else if (diameter > 0) row_radius = diameter / 2;
        else row_radius = float(rf) / 2;
    }
    if (col_radius <= 0) {
        if (col_diameter > 0) col_radius = col_diameter / 2;
        else if (radius > 0) col_radius = radius;
        else if (diameter > 0) col_radius = diameter / 2;
        else col_radius = float(cf) / 2;
    }
    const float row_center = float(rf) / 2, col_center = float(cf) / 2;
    const float row_radius_sq = (float)std::pow((double)row_radius, 2.0);
    const float col_radius_sq = (float)std::pow((double)col_radius, 2.0);
    const float sq_mult = row_radius_sq * col_radius_sq;
    const int n = c.convolutional ? 1 : c.to->size;
    for (int idx = 0; idx < n; ++idx)
        for (int kr = 0; kr < rf; ++kr)
            for (int kc = 0; kc < cf; ++kc) {
                const float dr = (float)(kr + 0.5 - row_center);
                const float dc = (float)(kc + 0.5 - col_center);
                const float term = (dr * dr) * col_radius_sq +
                                   (dc * dc) * row_radius_sq;
                if (invert == (term <= sq_mult))
                    w[(long)idx * ks + kr * cf + kc] = value;
            }
}

// initialize_weights (weight_matrix.cpp:979-1034).
void State::init_weights(ConnState &cs) {
    Connection &c = *cs.conn;
    const Props cfg = c.weight_config();
    float *w = cs.arr("weights").f();
    const int nw = c.num_weights;
    if (cfg.has("fraction")) {
        const float fr = cfg.get_float("fraction", 1.0f);
        if (fr < 0 || fr > 1.0f)
            throw Error("Weight config fraction must be between 0 and 1!");
    }
    const std::string type = cfg.get("type", "flat");
    const float fraction = cfg.get_float("fraction", 1.0f);
    if (type == "flat") {
        const float v = cfg.get_float("weight", 1.0f);
        fill_fraction(w, nw, fraction, [v] { return v; });
    } else if (type == "random") {
        const float hi = cfg.get_float("max weight", c.max_weight);
        const float lo = cfg.get_float("min weight", 0.0f);
        std::uniform_real_distribution<float> d(lo, hi);
        fill_fraction(w, nw, fraction, [&] { return d(rng()); });
    } else if (type == "gaussian" || type == "log normal") {
        const float mean = cfg.get_float("mean", 1.0f);
        const float sd = cfg.get_float("std dev", 0.3f);
        if (sd < 0) throw Error("Weight config std_dev must be positive!");
        const double mx = c.max_weight;
        if (sd == 0.0f) {
            fill_fraction(w, nw, fraction, [mean] { return mean; });
        } else if (type == "gaussian") {
            std::normal_distribution<double> d(mean, sd);
            for (int i = 0; i < nw; ++i)
                w[i] = (float)std::min(mx, std::max(0.0, d(rng())));
        } else {
            std::lognormal_distribution<double> d(mean, sd);
            for (int i = 0; i < nw; ++i)
                w[i] = (float)std::min(mx, std::max(0.0, d(rng())));
        }
    } else if (type == "power law") {
        const float ex = std::fabs(cfg.get_float("exponent", 1.5f));
        const float hi = cfg.get_float("max weight", c.max_weight);
        const float lo = cfg.get_float("min weight", 0.0f);
        const float ca = (float)std::pow((double)hi, 1.0 - ex);
        const float cb = (float)std::pow((double)std::max(lo, 0.00001f), 1.0 - ex);
        const float coeff = ca - cb;
        const float pe = (float)(1.0 / (1.0 - ex));
        std::uniform_real_distribution<float> d(0.0f, 1.0f);
        fill_fraction(w, nw, fraction, [&] {
            return (float)std::pow((double)(coeff * d(rng()) + cb), (double)pe);
        });
    } else if (type == "specified") {
        const std::string ws = cfg.get("weight string", "");
        if (ws.empty())
            throw Error("Missing weight string for specified weight config!");
        std::stringstream ss(ws);
        for (int i = 0; i < nw; ++i) {
            float v = 0.0f;
            if (!(ss >> v) && i + 1 < nw)
                throw Error("Insufficient number of weights specified!");
            w[i] = v;
        }
    } else {
        throw Error("Unrecognized weight config type: " + type);
    }

    const int id = cfg.get_int("id", 0);
    auto &cb = Callbacks::get();
    if (cfg.has("callback"))
        cb.find(cb.weight, cfg.get("callback"), "weight")(id, nw, w);
    if (cfg.has("indices callback")) {
        auto f = cb.find(cb.indices, cfg.get("indices callback"), "indices");
        compute_indices(cs);
        f(id, nw, w, cs.arr("from indices").i(), cs.arr("to indices").i(),
          cs.arr("used").i());
    }
    if (cfg.has("distance callback")) {
        auto f = cb.find(cb.distance, cfg.get("distance callback"),
                         "distance");
        const float fs = cfg.has("from spacing")
            ? cfg.get_float("from spacing", 1.0f)
            : c.from->config.get_float("neuron spacing", 1.0f);
        const float ts = cfg.has("to spacing")
            ? cfg.get_float("to spacing", 1.0f)
            : c.to->config.get_float("neuron spacing", 1.0f);
        compute_distances(cs, fs, ts, cfg.get_float("x offset", 0.0f),
                          cfg.get_float("y offset", 0.0f));
        f(id, nw, w, cs.arr("distances").f());
    }
    if (cfg.has("delay callback")) {
        auto f = cb.find(cb.delay, cfg.get("delay callback"), "delay");
        if (!cs.has("delays"))
            cs.arrays["delays"] =
                make_array("delays", Elem::I32, nw, (uint32_t)c.delay);
        f(id, nw, w, cs.arr("delays").i());
    }
    if (Props *m = cfg.get_child("circular mask")) circular_mask(*m, w, c);
    for (auto *m : cfg.get_child_array("circular mask")) circular_mask(*m, w, c);
    if (!cfg.get_bool("diagonal", true)) {
        switch (c.type) {
            case SG_FULLY_CONNECTED:
                if (c.from->size != c.to->size)
                    throw Error("Attempted to clear diagonal of non-square "
                                "weight matrix!");
                for (int i = 0; i < c.from->size; ++i)
                    w[(long)i * c.from->size + i] = 0.0f;
                break;
            case SG_SUBSET: {
                const int fs = c.subset.from_size();
                if (fs != c.subset.to_size())
                    throw Error("Attempted to clear diagonal of non-square "
                                "weight matrix!");
                for (int i = 0; i < fs; ++i) w[(long)i * fs + i] = 0.0f;
                break;
            }
            case SG_CONVERGENT:
                circular_mask(Props({{"radius", "0.5"}, {"invert", "true"}}),
                              w, c);
                break;
            default:
                break;
        }
    }
}

// IzhikevichAttributes::process_weight_matrix
// (izhikevich_attributes.cpp:692-757).
void State::process_iz_matrix(ConnState &cs) {
    Connection &c = *cs.conn;
    const int nw = c.num_weights;
    cs.baseline_conductance = std::stof(c.param("conductance", "1.0"));
    cs.learning_rate = std::stof(c.param("learning rate", "0.1"));
    cs.stp_flag = c.param("short term plasticity", "true") == "true";
    std::fill(cs.arr("short term R").host.begin(),
              cs.arr("short term R").host.end(), fbits(1.0f));
    // The reference compares the connection type with opcode values here,
    // so FULLY_CONNECTED (== ADD) starts at U_ADD, SUBSET (== SUB) at U_SUB
    // and every other type at 0.
    const float stp_w0 = c.type == SG_FULLY_CONNECTED ? 0.5f
                       : c.type == SG_SUBSET ? 0.2f : 0.0f;
    std::fill(cs.arr("short term w").host.begin(),
              cs.arr("short term w").host.end(), fbits(stp_w0));
    std::fill(cs.arr("weight derivatives").host.begin(),
              cs.arr("weight derivatives").host.end(), fbits((float)0.000001));
    std::fill(cs.arr("time since spike").host.begin(),
              cs.arr("time since spike").host.end(), 32u);

    if (c.param("myelinated", "false") == "true") {
        if (!cs.has("delays"))
            cs.arrays["delays"] =
                make_array("delays", Elem::I32, nw, (uint32_t)c.delay);
    } else if (c.config.has("random delay")) {
        const int mx = std::stoi(c.param("random delay", "0"));
        if (mx > 31)
            throw Error("Randomized axons cannot have delays greater than 31!");
        if (!cs.has("delays"))
            cs.arrays["delays"] = make_array("delays", Elem::I32, nw);
        std::uniform_int_distribution<int> d(0, mx);
        int32_t *dl = cs.arr("delays").i();
        for (int i = 0; i < nw; ++i) dl[i] = d(rng());
    } else if (!cs.has("delays")) {
        const float fs = std::stof(c.from->param("neuron spacing", "0.1"));
        const float ts = std::stof(c.to->param("neuron spacing", "0.1"));
        compute_distances(cs, fs, ts, std::stof(c.param("x offset", "0.0")),
                          std::stof(c.param("y offset", "0.0")));
        const bool cap = c.param("cap delay", "false") == "true";
        cs.arrays["delays"] = make_array("delays", Elem::I32, nw);
        int32_t *dl = cs.arr("delays").i();
        const float *dist = cs.arr("distances").f();
        for (int i = 0; i < nw; ++i) {
            const int d = c.delay + (int)(dist[i] / 0.15f);
            if (d > 31 && !cap)
                throw Error("BIT delays cannot be greater than 31!");
            dl[i] = std::min(31, d);
        }
    }
}

// ---------------------------------------------------------------------------
// iSTDP tables (izhikevich_attributes.cpp:454-472): src for
// powf(t,10)*1e-6*2.29*expf(-1.1 t), dst for powf(t,10)*1e-7*2.6*expf(-0.94 t).
void istdp_tables(float src[32], float dst[32]) {
    for (int t = 0; t < 32; ++t) {
        volatile int tv = t;
        const int tt = tv;
        src[t] = (float)((double)powf((float)tt, 10.0f) * 0.000001 * 2.29 *
                         (double)expf((float)(-1.1 * tt)));
        dst[t] = (float)((double)powf((float)tt, 10.0f) * 0.0000001 * 2.6 *
                         (double)expf((float)(-0.94 * tt)));
    }
}

// ---------------------------------------------------------------------------
// Host <-> device.

void *State::scratch(size_t bytes) {
    if (bytes > scratch_bytes_) {
        if (scratch_) check_hip(hipFree(scratch_), "hipFree");
        check_hip(hipMalloc(&scratch_, bytes), "hipMalloc scratch");
        scratch_bytes_ = bytes;
    }
    return scratch_;
}

static bool needs_device(const std::string &key) {
    static const char *host_only[] = {
        "delays", "distances", "used", "from row indices",
        "from column indices", "from indices", "to row indices",
        "to column indices", "to indices", "time since spike",
        "nonzero counts"};
    for (auto *k : host_only)
        if (key == k) return false;
    return true;
}

static size_t tiled_elems(int slots, int K) {
    return (size_t)((slots + 63) / 64) * 64 * (size_t)K;
}

void State::upload(Array &a) {
    if (a.host.empty()) return;
    const size_t want = a.tiled ? tiled_elems(a.rows, a.cols) : a.size();
    if (a.dev && a.dev_elems != want) {
        check_hip(hipFree(a.dev), "hipFree");
        a.dev = nullptr;
    }
    if (!a.dev) {
        check_hip(hipMalloc(&a.dev, want * 4), "hipMalloc");
        a.dev_elems = want;
    }
    if (!a.tiled) {
        check_hip(hipMemcpy(a.dev, a.host.data(), a.bytes(),
                            hipMemcpyHostToDevice), "hipMemcpy H2D");
        return;
    }
    void *tmp = scratch(a.bytes());
    check_hip(hipMemcpy(tmp, a.host.data(), a.bytes(), hipMemcpyHostToDevice),
              "hipMemcpy H2D");
    check_hip(sg_hip_tile32(tmp, a.dev, a.rows, a.cols, 1, 0, 1, nullptr),
              "tile");
    check_hip(hipDeviceSynchronize(), "tile sync");
}

void State::download(Array &a) {
    if (a.host.empty() || !a.dev) return;
    if (!a.tiled) {
        check_hip(hipMemcpy(a.host.data(), a.dev, a.bytes(),
                            hipMemcpyDeviceToHost), "hipMemcpy D2H");
        return;
    }
    void *tmp = scratch(a.bytes());
    check_hip(sg_hip_tile32(a.dev, tmp, a.rows, a.cols, 1, 0, 0, nullptr),
              "untile");
    check_hip(hipMemcpy(a.host.data(), tmp, a.bytes(), hipMemcpyDeviceToHost),
              "hipMemcpy D2H");
}

void State::ensure_device() {
    if (device_ready_) return;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0)
        throw Error("No HIP device available: this library runs the "
                    "simulation on an MI355X GPU");
    float tab[64];
    istdp_tables(tab, tab + 32);
    check_hip(hipMalloc((void **)&d_tables_, sizeof(tab)), "hipMalloc");
    check_hip(hipMemcpy(d_tables_, tab, sizeof(tab), hipMemcpyHostToDevice),
              "hipMemcpy");
    device_ready_ = true;
}

// Per-destination-slot arrays ([slot][k] on the host) are tiled by 64
// destinations on the device; shared kernels and one-to-one arrays are
// copied as they are.
static bool conn_tiled(const ConnState &cs) {
    return !cs.conn->convolutional && cs.conn->type != SG_ONE_TO_ONE;
}

static void set_conn_layout(ConnState &cs, Array &a) {
    a.tiled = conn_tiled(cs) &&
              (long)cs.num_dest * cs.k_per_dest == (long)a.size();
    a.rows = cs.num_dest;
    a.cols = cs.k_per_dest;
}

void State::upload_conn_derived(ConnState &cs) {
    Connection &c = *cs.conn;
    const bool conductance = cs.iz && !cs.direct && cs.iz_kind != SG_IZ_GAP;
    // Base activators reading a BIT source get a spike-plane scratch buffer
    // (one bit per source neuron).
    if (c.type == SG_SPARSE) {
        // source index per (slot, k) in the synapse layout, -1 past each
        // slot's nonzero count
        const int K = cs.k_per_dest;
        std::vector<int32_t> idx(tiled_elems(cs.num_dest, K), -1);
        const int32_t *cnt = cs.arr("nonzero counts").i();
        const int32_t *fi = cs.arr("from indices").i();
        for (int slot = 0; slot < cs.num_dest; ++slot) {
            if (cnt[slot] < 0 || cnt[slot] > K)
                throw Error("Invalid nonzero count in " + c.str());
            for (int k = 0; k < cnt[slot]; ++k) {
                const int32_t f = fi[(long)slot * K + k];
                // a source outside the layer would be read out of bounds
                // (e.g. the reference's randomized convergent projection
                // without wrap, whose index row * rows + column can leave
                // a non-square layer, weight_matrix.cpp:436-437)
                if (f < 0 || f >= c.from->size)
                    throw Error("Error in " + c.str() + ":\n  sparse source "
                                "index " + std::to_string(f) +
                                " outside the source layer");
                idx[(((long)slot >> 6) * K + k) * 64 + (slot & 63)] = f;
            }
        }
        if (cs.d_sparse_from) check_hip(hipFree(cs.d_sparse_from), "hipFree");
        check_hip(hipMalloc((void **)&cs.d_sparse_from, idx.size() * 4),
                  "hipMalloc sparse");
        check_hip(hipMemcpy(cs.d_sparse_from, idx.data(), idx.size() * 4,
                            hipMemcpyHostToDevice), "hipMemcpy sparse");
        // per 64-destination tile: the fewest and the most nonzero visits
        // of its slots
        const int tiles = (cs.num_dest + 63) / 64;
        std::vector<int32_t> tc((size_t)tiles * 2);
        for (int t = 0; t < tiles; ++t) {
            int lo = K, hi = 0;
            for (int slot = t * 64; slot < std::min(cs.num_dest, t * 64 + 64);
                 ++slot) {
                lo = std::min(lo, cnt[slot]);
                hi = std::max(hi, cnt[slot]);
            }
            tc[(size_t)t * 2] = lo;
            tc[(size_t)t * 2 + 1] = hi;
        }
        if (cs.d_tile_counts) check_hip(hipFree(cs.d_tile_counts), "hipFree");
        check_hip(hipMalloc((void **)&cs.d_tile_counts, tc.size() * 4),
                  "hipMalloc tile counts");
        check_hip(hipMemcpy(cs.d_tile_counts, tc.data(), tc.size() * 4,
                            hipMemcpyHostToDevice), "hipMemcpy tile counts");
    }
    if ((!cs.iz || cs.direct) && cs.from->out_type == SG_OUT_BIT &&
        !cs.d_plane)
        check_hip(hipMalloc((void **)&cs.d_plane,
                            (size_t)(c.from->size + 63) / 64 * 8),
                  "hipMalloc plane");
    if (!cs.iz) return;
    const int from_n = c.from->size;
    Array &tss = cs.arr("time since spike");
    std::vector<int32_t> per_source(from_n, 32);
    const size_t nn = std::min((size_t)from_n, tss.size());
    std::memcpy(per_source.data(), tss.host.data(), nn * 4);
    for (int b = 0; b < 2; ++b) {
        if (!cs.d_tss[b])
            check_hip(hipMalloc((void **)&cs.d_tss[b], (size_t)from_n * 4),
                      "hipMalloc tss");
    }
    check_hip(hipMemcpy(cs.d_tss[0], per_source.data(), (size_t)from_n * 4,
                        hipMemcpyHostToDevice), "hipMemcpy tss");
    cs.tss_cur = 0;
    if (!conductance) return;

    // Delays in the synapse layout, and per source the delays of its last
    // (up to) 32 visits in the reference's serial (destination-major) order,
    // newest first, for the time-since-spike pass.
    std::vector<int32_t> visits(from_n, 0);
    enumerate_synapses(cs, [&](int, int, int, int r, int col, int, int,
                               long) {
        ++visits[r * c.from->cols + col];
    });
    const bool tiled = conn_tiled(cs);
    // Tiled types: the record's visit-group-major order (see sg_hip.h).
    const long tiles = (cs.num_dest + 63) / 64;
    const size_t nw = tiled ? tiled_elems(cs.num_dest,
                                          (cs.k_per_dest + 7) / 8 * 8)
                            : (size_t)c.num_weights;
    std::vector<uint8_t> dly(nw, 0);
    std::vector<uint8_t> rev((size_t)from_n * 32, 0);
    std::vector<int32_t> seen(from_n, 0);
    const int32_t *dl = cs.arr("delays").i();
    enumerate_synapses(cs, [&](int slot, int, int k, int r, int col, int,
                               int, long wi) {
        const int f = r * c.from->cols + col;
        const int j = ++seen[f];           // 1-based position of this visit
        const int back = visits[f] - j;    // visits of f after this one
        const uint8_t d = (uint8_t)(dl[wi] & 31);
        if (back < 32) rev[(size_t)f * 32 + back] = d;
        const long di = tiled
            ? (((long)(k >> 3) * tiles + (slot >> 6)) * 8 + (k & 7)) * 64 +
                  (slot & 63)
            : wi;
        dly[di] = d;
    });
    std::vector<uint8_t> vcap(from_n);
    for (int f = 0; f < from_n; ++f) vcap[f] = (uint8_t)std::min(32, visits[f]);
    auto put = [&](uint8_t *&dst, const std::vector<uint8_t> &v) {
        if (dst) check_hip(hipFree(dst), "hipFree");
        check_hip(hipMalloc((void **)&dst, v.size()), "hipMalloc");
        check_hip(hipMemcpy(dst, v.data(), v.size(), hipMemcpyHostToDevice),
                  "hipMemcpy");
    };
    put(cs.d_delays, dly);
    put(cs.d_rev_delays, rev);
    put(cs.d_visits, vcap);
}

// The seven per-synapse float arrays of a tiled Izhikevich conductance
// connection, in the field order of its interleaved device record.
static const char *kRecordFields[SG_IZ_FIELDS] = {
    "weights", "short term R", "short term w", "presyn trace",
    "weight derivatives", "short trace", "long trace"};

static bool record_packed(const ConnState &cs) {
    return cs.iz && !cs.direct && cs.iz_kind >= 0 && cs.iz_kind != SG_IZ_GAP &&
           conn_tiled(cs);
}

// Compact record fields: the arrays they hold (DW: see compact_dw).
static const char *kCompactFields[SG_IZC_FIELDS] = {
    "weights", "weight derivatives", "short trace", "long trace"};

static bool active_weight(float w) { return w >= 0.0001f; }

// Per source {R, w, presynaptic trace, 0} for a compact record: true when
// every visited synapse of weight >= 1e-4 from one source carries the same
// three values (bit for bit).  Sources without such a synapse keep the
// first visited synapse's values (no active synapse reads them).
static bool compact_sources(const ConnState &cs, std::vector<float> &out) {
    const Connection &c = *cs.conn;
    const int n = c.from->size;
    out.assign((size_t)n * 4, 0.0f);
    std::vector<uint8_t> seen(n, 0);
    const float *w = const_cast<ConnState &>(cs).arr("weights").f();
    const float *R = const_cast<ConnState &>(cs).arr("short term R").f();
    const float *W = const_cast<ConnState &>(cs).arr("short term w").f();
    const float *P = const_cast<ConnState &>(cs).arr("presyn trace").f();
    bool ok = true;
    enumerate_synapses(cs, [&](int, int, int, int r, int col, int, int,
                               long wi) {
        if (!ok) return;
        const int f = r * c.from->cols + col;
        const bool act = active_weight(w[wi]);
        float *o = &out[(size_t)f * 4];
        if (seen[f] < 2 && (act || !seen[f])) {
            o[0] = R[wi];
            o[1] = W[wi];
            o[2] = P[wi];
            seen[f] = act ? 2 : 1;
            return;
        }
        if (act && (fbits(o[0]) != fbits(R[wi]) || fbits(o[1]) != fbits(W[wi]) ||
                    fbits(o[2]) != fbits(P[wi])))
            ok = false;
    });
    return ok;
}

// DW field of a compact record in host layout: the derivative of an active
// synapse, R * w of a frozen one (weight below 1e-4: the updater never
// touches it, and its activation weighs w * (R * w_stp)).
static std::vector<float> compact_dw(const ConnState &cs) {
    ConnState &m = const_cast<ConnState &>(cs);
    const float *w = m.arr("weights").f();
    const float *R = m.arr("short term R").f();
    const float *W = m.arr("short term w").f();
    const float *D = m.arr("weight derivatives").f();
    std::vector<float> out(D, D + m.arr("weight derivatives").size());
    enumerate_synapses(cs, [&](int, int, int, int, int, int, int, long wi) {
        if (!active_weight(w[wi])) out[wi] = R[wi] * W[wi];
    });
    return out;
}

void State::to_device(bool allow_compact) {
    ensure_device();
    if (!on_host_) return;
    ++device_epoch_;
    for (auto &ls : layers_) {
        upload(ls->input);
        upload(ls->output);
        for (auto &kv : ls->vars) upload(kv.second);
        if (ls->model == Model::IZHIKEVICH && !ls->d_plane)
            check_hip(hipMalloc((void **)&ls->d_plane,
                                (size_t)(ls->layer->size + 63) / 64 * 8),
                      "hipMalloc layer plane");
    }
    for (auto &cs : conns_) {
        cs->packed = record_packed(*cs);
        cs->compact = false;
        if (cs->packed) {
            std::vector<float> src;
            cs->compact = allow_compact && compact_sources(*cs, src);
            const int nf = cs->compact ? (int)SG_IZC_FIELDS : (int)SG_IZ_FIELDS;
            const size_t n = tiled_elems(cs->num_dest,
                                         (cs->k_per_dest + 7) / 8 * 8) * nf;
            if (cs->d_record && cs->record_fields != nf) {
                check_hip(hipFree(cs->d_record), "hipFree record");
                cs->d_record = nullptr;
            }
            if (!cs->d_record)
                check_hip(hipMalloc((void **)&cs->d_record, n * 4),
                          "hipMalloc record");
            cs->record_fields = nf;
            std::vector<float> dw;
            if (cs->compact) dw = compact_dw(*cs);
            for (int f = 0; f < nf; ++f) {
                Array &a = cs->arr(cs->compact ? kCompactFields[f]
                                               : kRecordFields[f]);
                const void *h = (cs->compact && f == SG_IZC_DW)
                    ? (const void *)dw.data() : (const void *)a.host.data();
                void *tmp = scratch(a.bytes());
                check_hip(hipMemcpy(tmp, h, a.bytes(), hipMemcpyHostToDevice),
                          "hipMemcpy H2D");
                check_hip(sg_hip_record32(tmp, cs->d_record, cs->num_dest,
                                          cs->k_per_dest, f, nf, 1, nullptr),
                          "record upload");
                check_hip(hipDeviceSynchronize(), "tile sync");
            }
            if (cs->compact) {
                if (!cs->d_src_stp)
                    check_hip(hipMalloc((void **)&cs->d_src_stp,
                                        src.size() * 4),
                              "hipMalloc source stp");
                check_hip(hipMemcpy(cs->d_src_stp, src.data(), src.size() * 4,
                                    hipMemcpyHostToDevice),
                          "hipMemcpy source stp");
            }
        }
        for (auto &kv : cs->arrays) {
            if (!needs_device(kv.first)) continue;
            if (cs->packed && kv.first != "weights" &&
                std::find(std::begin(kRecordFields), std::end(kRecordFields),
                          kv.first) != std::end(kRecordFields))
                continue;
            set_conn_layout(*cs, kv.second);
            if (cs->packed && kv.first == "weights") continue;
            upload(kv.second);
        }
        upload_conn_derived(*cs);
        const float *w = cs->arr("weights").f();
        bool fin = true;
        for (size_t i = 0, n = cs->arr("weights").size(); i < n && fin; ++i)
            fin = std::isfinite(w[i]);
        cs->finite_weights = fin;
    }
    check_hip(hipDeviceSynchronize(), "upload sync");
    on_host_ = false;
}

// A compact record back to the host arrays: weights and traces from the
// record; for every visited synapse of weight >= 1e-4 its derivative from
// the DW field and R, w and the presynaptic trace from its source.  A
// frozen synapse's derivative, R, w and trace never changed: the host
// copies still hold them.
void State::download_compact(ConnState &cs) {
    std::vector<float> dw(cs.arr("weight derivatives").size());
    for (int f = 0; f < SG_IZC_FIELDS; ++f) {
        Array &a = cs.arr(kCompactFields[f]);
        void *dst = f == SG_IZC_DW ? (void *)dw.data() : (void *)a.host.data();
        void *tmp = scratch(a.bytes());
        check_hip(sg_hip_record32(cs.d_record, tmp, cs.num_dest, cs.k_per_dest,
                                  f, SG_IZC_FIELDS, 0, nullptr),
                  "record download");
        check_hip(hipMemcpy(dst, tmp, a.bytes(), hipMemcpyDeviceToHost),
                  "hipMemcpy D2H");
    }
    const Connection &c = *cs.conn;
    std::vector<float> src((size_t)c.from->size * 4);
    check_hip(hipMemcpy(src.data(), cs.d_src_stp, src.size() * 4,
                        hipMemcpyDeviceToHost), "hipMemcpy source stp");
    const float *w = cs.arr("weights").f();
    float *R = cs.arr("short term R").f();
    float *W = cs.arr("short term w").f();
    float *P = cs.arr("presyn trace").f();
    float *D = cs.arr("weight derivatives").f();
    enumerate_synapses(cs, [&](int, int, int, int r, int col, int, int,
                               long wi) {
        if (!active_weight(w[wi])) return;
        const float *o = &src[(size_t)(r * c.from->cols + col) * 4];
        R[wi] = o[0];
        W[wi] = o[1];
        P[wi] = o[2];
        D[wi] = dw[wi];
    });
}

void State::to_host() {
    if (on_host_) return;
    flush_pending_updates(*this);
    check_hip(hipDeviceSynchronize(), "download sync");
    for (auto &ls : layers_) {
        download(ls->input);
        download(ls->output);
        for (auto &kv : ls->vars) download(kv.second);
    }
    for (auto &cs : conns_) {
        if (cs->packed && cs->compact) {
            download_compact(*cs);
        } else if (cs->packed) {
            for (int f = 0; f < SG_IZ_FIELDS; ++f) {
                Array &a = cs->arr(kRecordFields[f]);
                void *tmp = scratch(a.bytes());
                check_hip(sg_hip_record32(cs->d_record, tmp, cs->num_dest,
                                          cs->k_per_dest, f, SG_IZ_FIELDS, 0,
                                          nullptr),
                          "record download");
                check_hip(hipMemcpy(a.host.data(), tmp, a.bytes(),
                                    hipMemcpyDeviceToHost), "hipMemcpy D2H");
            }
        } else {
            for (auto &kv : cs->arrays)
                if (needs_device(kv.first)) download(kv.second);
        }
        if (cs->iz) {
            Array &tss = cs->arr("time since spike");
            const size_t nn = std::min((size_t)cs->conn->from->size,
                                       tss.size());
            check_hip(hipMemcpy(tss.host.data(), cs->d_tss[cs->tss_cur],
                                nn * 4, hipMemcpyDeviceToHost),
                      "hipMemcpy tss");
        }
    }
    on_host_ = true;
}

Array *State::neuron_data(Layer *layer, const std::string &key) {
    to_host();
    LayerState *ls = layer_state(layer);
    if (key == "input") return &ls->input;
    if (key == "output") return &ls->output;
    auto it = ls->vars.find(key);
    if (it == ls->vars.end())
        throw Error("Failed to retrieve neuron data \"" + key + "\"");
    return &it->second;
}

Array *State::weight_matrix(Connection *conn, const std::string &key) {
    to_host();
    ConnState *cs = conn_state(conn);
    auto it = cs->arrays.find(key);
    if (it == cs->arrays.end())
        throw Error("Failed to retrieve data \"" + key +
                    "\" in WeightMatrix for connection:" + conn->str());
    return &it->second;
}

size_t State::network_bytes() const {
    size_t b = 0;
    for (auto &ls : layers_) {
        b += ls->input.bytes() + ls->output.bytes();
        for (auto &kv : ls->vars) b += kv.second.bytes();
    }
    for (auto &cs : conns_)
        for (auto &kv : cs->arrays) b += kv.second.bytes();
    return b;
}

// ---------------------------------------------------------------------------
// Checkpoint file: a sequence of (PointerKey{id, hash(key), bytes}, bytes)
// records in (id, hash) order (state/state.cpp:365-447).

namespace {
struct Key {
    size_t hash, type, bytes;
};
}

void State::save(const std::string &path) {
    to_host();
    std::map<std::pair<size_t, size_t>, Array *> recs;
    std::hash<std::string> h;
    for (auto &ls : layers_) {
        const size_t id = ls->layer->id;
        recs[{id, h("input")}] = &ls->input;
        recs[{id, h("output")}] = &ls->output;
        for (auto &kv : ls->vars) recs[{id, h(kv.first)}] = &kv.second;
    }
    for (auto &cs : conns_)
        for (auto &kv : cs->arrays)
            recs[{cs->conn->id, h(kv.first)}] = &kv.second;
    std::ofstream out(path, std::ios::binary);
    if (!out) throw Error("Could not open file: " + path);
    for (auto &r : recs) {
        Key k{r.first.first, r.first.second, r.second->bytes()};
        if (k.bytes == 0) continue;
        if (!out.write((const char *)&k, sizeof(k)) ||
            !out.write((const char *)r.second->host.data(), k.bytes))
            throw Error("Error writing state to file!");
    }
}

void State::load(const std::string &path) {
    std::ifstream in(path, std::ios::binary);
    if (!in) throw Error("Could not open file: " + path);
    to_host();
    std::map<std::pair<size_t, size_t>, Array *> recs;
    std::hash<std::string> h;
    for (auto &ls : layers_) {
        const size_t id = ls->layer->id;
        recs[{id, h("input")}] = &ls->input;
        recs[{id, h("output")}] = &ls->output;
        for (auto &kv : ls->vars) recs[{id, h(kv.first)}] = &kv.second;
    }
    for (auto &cs : conns_)
        for (auto &kv : cs->arrays)
            recs[{cs->conn->id, h(kv.first)}] = &kv.second;
    Key k;
    while (in.read((char *)&k, sizeof(k))) {
        auto it = recs.find({k.hash, k.type});
        if (it == recs.end() || k.bytes % 4 != 0) {
            in.ignore(k.bytes);
            continue;
        }
        Array *a = it->second;
        if (a->bytes() != k.bytes) a->host.assign(k.bytes / 4, 0);
        if (!in.read((char *)a->host.data(), k.bytes))
            throw Error("Error reading data from file!");
    }
}

}  // namespace sg
