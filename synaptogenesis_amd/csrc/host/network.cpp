#include "network.h"

#include <functional>

namespace sg {

void add_dendrites(Layer &l, Dendrite *parent,
                   const std::vector<Props *> &configs);

int connection_type_from(const std::string &s) {   // util/constants.h:88-94
    if (s == "fully connected") return SG_FULLY_CONNECTED;
    if (s == "subset") return SG_SUBSET;
    if (s == "one to one") return SG_ONE_TO_ONE;
    if (s == "convergent") return SG_CONVERGENT;
    if (s == "divergent") return SG_DIVERGENT;
    throw Error("Unrecognized ConnectionType: " + s);
}

int opcode_from(const std::string &s) {           // util/constants.h:137-149
    static const char *names[] = {"add", "sub", "mult", "div", "pool",
                                  "reward", "modulate", "gap",
                                  "add_heaviside", "sub_heaviside",
                                  "mult_heaviside"};
    for (int i = 0; i < 11; ++i)
        if (s == names[i]) return i;
    throw Error("Unrecognized Opcode: " + s);
}

static ClusterType cluster_from(const std::string &s) {
    if (s == "parallel") return CLUSTER_PARALLEL;
    if (s == "sequential") return CLUSTER_SEQUENTIAL;
    if (s == "feedforward") return CLUSTER_FEEDFORWARD;
    throw Error("Unrecognized ClusterType: " + s);
}

int Layer::max_out_delay() const {                  // layer.cpp get_max_delay
    int d = 0;
    for (auto *c : outputs) d = std::max(d, c->delay);
    return d;
}

std::string Layer::str() const {
    return "[Layer: " + name + " (" + structure->name + ")]";
}

std::string Connection::str() const {               // connection.cpp:251-259
    return "[Connection: " + from->name + " (" + from->structure->name +
           ") -> " + to->name + " (" + to->structure->name + ") " +
           "{Node: " + node + "} name=" + name + "]";
}

int Connection::matrix_rows() const {
    if (convolutional) return 1;
    if (type == SG_SUBSET) return subset.to_size();
    return to->size;
}

int Connection::matrix_columns() const {
    if (type == SG_SUBSET) return subset.from_size();
    if (type == SG_SPARSE) return num_weights / to->size;
    if (type == SG_CONVERGENT || type == SG_DIVERGENT) return arbor.total();
    if (type == SG_ONE_TO_ONE) return 1;
    return from->size;
}

Layer *Structure::layer(const std::string &n) const {
    for (auto &l : layers)
        if (l->name == n) return l.get();
    return nullptr;
}

static ArborConfig parse_arbor(const Props &c) {
    ArborConfig a;
    a.row_field = c.get_int("row field size", 0);
    a.col_field = c.get_int("column field size", 0);
    a.row_stride = c.get_int("row stride", 1);
    a.col_stride = c.get_int("column stride", 1);
    a.row_spacing = c.get_int("row spacing", 1);
    a.col_spacing = c.get_int("column spacing", 1);
    a.row_offset = c.get_int("row offset", 0);
    // The column offset is read from the "row offset" key
    // (connection_config.cpp:16).
    a.col_offset = c.get_int("row offset", 0);
    a.wrap = c.get_bool("wrap", false);
    if (c.has("field size"))
        a.row_field = a.col_field = c.get_int("field size", 0);
    if (c.has("stride"))
        a.row_stride = a.col_stride = c.get_int("stride", 0);
    if (c.has("spacing"))
        a.row_spacing = a.col_spacing = c.get_int("spacing", 0);
    if (c.has("offset"))
        a.row_offset = a.col_offset = c.get_int("offset", 0);
    if (a.row_field <= 0 || a.col_field <= 0)
        throw Error("Unspecified field size for arborized config!");
    if (a.row_stride < 0 || a.col_stride < 0)
        throw Error("Cannot have negative stride in arborized config!");
    if (a.row_spacing < 1 || a.col_spacing < 1)
        throw Error("Cannot have zero or negative spacing in arborized config!");
    if (!c.has("row offset") && !c.has("offset"))
        a.row_offset = -a.row_field / 2;
    if (!c.has("column offset") && !c.has("offset"))
        a.col_offset = -a.col_field / 2;
    return a;
}

static SubsetConfig parse_subset(const Props &c) {
    SubsetConfig s;
    s.from_row_start = c.get_int("from row start", 0);
    s.from_row_end = c.get_int("from row end", 0);
    s.to_row_start = c.get_int("to row start", 0);
    s.to_row_end = c.get_int("to row end", 0);
    s.from_col_start = c.get_int("from column start", 0);
    s.from_col_end = c.get_int("from column end", 0);
    s.to_col_start = c.get_int("to column start", 0);
    s.to_col_end = c.get_int("to column end", 0);
    if (s.from_row_end - s.from_row_start <= 0 ||
        s.to_row_end - s.to_row_start <= 0 ||
        s.from_col_end - s.from_col_start <= 0 ||
        s.to_col_end - s.to_col_start <= 0)
        throw Error("Invalid subset dimensions!");
    return s;
}

Network::Network(const Props &config) : config_(config) {
    for (auto *s : config.get_child_array("structures")) add_structure(*s);
    for (auto *c : config.get_child_array("connections")) add_connection(*c);
}

void Network::add_structure(const Props &sc) {
    if (!sc.has("name")) throw Error("Unspecified name for structure!");
    std::unique_ptr<Structure> st(new Structure());
    st->name = sc.get("name");
    st->cluster = cluster_from(sc.get("type", "parallel"));
    for (auto &o : structures_)
        if (o->name == st->name)
            throw Error("Repeated structure name: " + st->name);
    for (auto *lc : sc.get_child_array("layers")) {
        if (!lc->has("name"))
            throw Error("Attempted to construct LayerConfig without name!");
        if (!lc->has("neural model"))
            throw Error("Attempted to construct LayerConfig without neural "
                        "model!");
        if (lc->has("image"))
            throw Error("Image-sized layers are not supported");
        std::unique_ptr<Layer> l(new Layer());
        l->structure = st.get();
        l->config = *lc;
        l->name = lc->get("name");
        l->neural_model = lc->get("neural model");
        l->rows = lc->get_int("rows", 0);
        l->cols = lc->get_int("columns", 0);
        l->size = l->rows * l->cols;
        l->plastic = lc->get_bool("plastic", false);
        l->id = std::hash<std::string>()(st->name + "/" + l->name);
        if (l->rows <= 0 || l->cols <= 0)
            throw Error("Invalid dimensions for layer " + l->name);
        l->dendrites.emplace_back(new Dendrite());
        l->dendrites[0]->name = "root";
        add_dendrites(*l, l->dendrites[0].get(),
                      lc->get_child_array("dendrites"));
        l->registers = (int)l->dendrites.size();
        if (st->layer(l->name))
            throw Error("Repeated layer name : \"" + l->name + "\"");
        st->layers.push_back(std::move(l));
    }
    structures_.push_back(std::move(st));
}

// Layer::add_dendrites (layer.cpp:135-152): children in config order,
// each followed by its own subtree; a new node takes the next register.
void add_dendrites(Layer &l, Dendrite *parent,
                   const std::vector<Props *> &configs) {
    for (Props *dc : configs) {
        if (!dc->has("name"))
            throw Error("Attempted to dendrite without name to layer!");
        const std::string name = dc->get("name");
        if (parent->second_order)      // dendritic_node.cpp:81-86
            throw Error("Error in dendritic node " + name + " of " + l.str() +
                        "\n  Second order dendritic nodes cannot have "
                        "internal children!");
        if (l.dendrite(name))
            throw Error("Error in dendritic node of " + l.str() + "\n"
                        "  Duplicate dendritic node name: " + name + "!");
        std::unique_ptr<Dendrite> d(new Dendrite());
        d->name = name;
        d->opcode = opcode_from(dc->get("opcode", "add"));
        d->init = dc->get_float("init", 0.0f);
        d->second_order = dc->get_bool("second order", false);
        d->reg = (int)l.dendrites.size();
        Dendrite *child = d.get();
        parent->children.push_back(child);
        l.dendrites.push_back(std::move(d));
        add_dendrites(l, child, dc->get_child_array("children"));
    }
}

Dendrite *Layer::dendrite(const std::string &n) const {
    for (auto &d : dendrites)
        if (d->name == n) return d.get();
    return nullptr;
}

void Network::add_connection(const Props &cc) {
    if (!cc.has("from layer"))
        throw Error("Attempted to construct ConnectionConfig without source "
                    "layer!");
    if (!cc.has("to layer"))
        throw Error("Attempted to construct ConnectionConfig without "
                    "destination layer!");
    const std::string fl = cc.get("from layer"), tl = cc.get("to layer");
    // network.cpp:152-217: with one structure names may be omitted;
    // otherwise the layer name is searched in every structure.
    auto find = [&](const std::string &key, const std::string &lname) {
        if (structures_.size() == 1) return structures_[0].get();
        if (cc.has(key)) {
            Structure *s = structure(cc.get(key));
            if (!s) throw Error("Could not find structure: " + cc.get(key));
            return s;
        }
        Structure *found = nullptr;
        for (auto &s : structures_)
            if (s->layer(lname)) {
                if (found) throw Error("Ambiguous layer " + lname);
                found = s.get();
            }
        if (!found) throw Error("Could not find layer " + lname);
        return found;
    };
    Structure *fs = find("from structure", fl);
    Structure *ts = find("to structure", tl);
    Layer *from = fs->layer(fl), *to = ts->layer(tl);
    if (!from || !to)
        throw Error("Could not find layer for connection " + fl + " -> " + tl);

    std::unique_ptr<Connection> c(new Connection());
    c->config = cc;
    c->name = cc.get("name", "");
    c->from = from;
    c->to = to;
    c->plastic = cc.get_bool("plastic", true);
    c->delay = cc.get_int("delay", 0);
    c->max_weight = cc.get_float("max weight", 1.0f);
    c->type = connection_type_from(cc.get("type", "fully connected"));
    c->opcode = opcode_from(cc.get("opcode", "add"));
    c->sparse = cc.get_bool("sparse", false);
    c->convolutional = cc.get_bool("convolutional", false);
    c->recurrent = cc.get_bool("recurrent", false);
    if (c->delay < 0)
        throw Error("Attempted to construct ConnectionConfig with negative "
                    "delay!");
    // A randomized projection is sparse (connection.cpp:138-139); its
    // sources are drawn when the weights are built
    // (weight_matrix.cpp:356-412).
    c->randomized = cc.get_bool("randomized projection", false);
    if (c->randomized) {
        c->sparse = true;
        if (c->convolutional)
            throw Error("Error intializing weight matrix for " + c->str() +
                        "\n   Cannot randomize projection on convolutional "
                        "weight matrix!");
        if (c->type == SG_FULLY_CONNECTED || c->type == SG_SUBSET)
            throw Error("Error intializing weight matrix for " + c->str() +
                        "\n   Cannot randomize projection on fully "
                        "connected or subset weight matrix!");
        // the reference builds a divergent one by inverting a randomized
        // convergent matrix into slots it never bounds-checks
        // (weight_matrix.cpp:444-503)
        if (c->type == SG_DIVERGENT)
            throw Error("Randomized divergent projections are not supported "
                        "(" + c->str() + ")");
    }
    if (c->sparse && c->convolutional)   // connection_config.cpp:234-236
        throw Error("Convolutional connections cannot be sparse!");
    if (c->sparse && c->type == SG_SUBSET)
        throw Error("Sparse SUBSET connections are not supported (the "
                    "sparse matrix indexes subset rows as layer "
                    "destinations)");
    if (c->convolutional && c->type != SG_CONVERGENT &&
        c->type != SG_DIVERGENT)
        throw Error("Only convergent/divergent connections can be "
                    "convolutional!");
    switch (c->type) {
        case SG_FULLY_CONNECTED:
            c->num_weights = from->size * to->size;
            break;
        case SG_SUBSET: {
            Props *sub = cc.get_child("subset config");
            if (!sub)
                throw Error("Attempted to create SUBSET connection without "
                            "specifying subset configuration!");
            c->subset = parse_subset(*sub);
            const SubsetConfig &s = c->subset;
            if (s.from_row_start < 0 || s.from_col_start < 0 ||
                s.to_row_start < 0 || s.to_col_start < 0 ||
                s.from_row_end > from->rows || s.from_col_end > from->cols ||
                s.to_row_end > to->rows || s.to_col_end > to->cols)
                throw Error("Invalid connection config!");
            c->num_weights = s.from_size() * s.to_size();
            break;
        }
        case SG_ONE_TO_ONE:
            if (from->rows != to->rows || from->cols != to->cols)
                throw Error("Cannot connect differently sized layers "
                            "one-to-one!");
            c->num_weights = to->size;
            break;
        case SG_CONVERGENT:
        case SG_DIVERGENT: {
            Props *ar = cc.get_child("arborized config");
            if (!ar)
                throw Error("Attempted to create arborized connection "
                            "without specifying arborized configuration!");
            c->arbor = parse_arbor(*ar);
            if (c->type == SG_DIVERGENT &&
                (c->arbor.row_stride == 0 || c->arbor.col_stride == 0))
                throw Error("Divergent connections cannot have zero stride!");
            c->num_weights = c->convolutional
                ? c->arbor.total() : to->size * c->arbor.total();
            break;
        }
    }
    c->dense_type = c->type;
    c->dense_num_weights = c->num_weights;
    const std::string dname = cc.get("dendrite", "root");
    Dendrite *node = to->dendrite(dname);
    if (!node)
        throw Error("Error in " + to->str() + ":\n"
                    "  Cannot find Dendritic Node " + dname + "!");
    c->node = node->name;
    c->reg = node->reg;
    // connection.cpp:141-190: second-order roles and their restrictions
    if (node->second_order) {
        c->so_host = node->host == nullptr;
        c->so_slave = !c->so_host;
        if (c->plastic)
            throw Error("Error in " + c->str() + ":\n  Plastic second "
                        "order connections are not supported!");
        if (c->sparse)
            throw Error("Error in " + c->str() + ":\n  Sparse second "
                        "order connections are not supported!");
        if (c->so_slave &&
            (c->type != node->host->type ||
             c->convolutional != node->host->convolutional ||
             c->num_weights != node->host->num_weights))
            throw Error("Error in " + c->str() + ":\n  Second order "
                        "connection does not match host connection!");
        // A convolutional slave folds like a one-to-one connection over its
        // source layer (kernel.cpp:223-229), so its field covers that layer
        // with one visit per source.
        if (c->so_slave && c->convolutional &&
            (c->arbor.total() != from->size || c->arbor.row_spacing != 1 ||
             c->arbor.col_spacing != 1 || c->arbor.row_stride != 0 ||
             c->arbor.col_stride != 0))
            throw Error("Error in " + c->str() + ":\n  Second order "
                        "convolutional connections must have fields that "
                        "are the size of the input layer, and must have 0 "
                        "stride and 1 spacing!");
    }
    if (c->so_host) node->host = c.get();
    else node->conns.push_back(c.get());
    c->id = std::hash<std::string>()(c->str());
    from->outputs.push_back(c.get());
    to->inputs.push_back(c.get());
    ts->connections.push_back(std::move(c));
}

std::vector<Layer *> Network::layers() const {
    std::vector<Layer *> out;
    for (auto &s : structures_)
        for (auto &l : s->layers) out.push_back(l.get());
    return out;
}

std::vector<Connection *> Network::connections() const {
    std::vector<Connection *> out;
    for (auto &s : structures_)
        for (auto &c : s->connections) out.push_back(c.get());
    return out;
}

Structure *Network::structure(const std::string &name) const {
    for (auto &s : structures_)
        if (s->name == name) return s.get();
    return nullptr;
}

Layer *Network::layer(const std::string &name) const {
    Layer *found = nullptr;
    for (auto *l : layers())
        if (l->name == name) found = l;
    return found;
}

Connection *Network::connection(const std::string &name) const {
    Connection *found = nullptr;
    for (auto *c : connections())
        if (c->name == name) found = c;
    return found;
}

int Network::max_layer_size() const {
    int m = 0;
    for (auto *l : layers()) m = std::max(m, l->size);
    return m;
}

}  // namespace sg
