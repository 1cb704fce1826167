// Network description: structures, layers and connections resolved from a
// Props tree, with the reference's defaults and validation
// (network/{network,structure,layer,connection,connection_config}.cpp).
#ifndef SG_NETWORK_H
#define SG_NETWORK_H

#include <memory>
#include <string>
#include <vector>

#include "props.h"
#include "sg_hip.h"

namespace sg {

enum ClusterType { CLUSTER_PARALLEL = 0, CLUSTER_SEQUENTIAL = 1,
                   CLUSTER_FEEDFORWARD = 2 };

struct Structure;
struct Connection;

struct ArborConfig {   // network/connection_config.cpp:8-45
    int row_field = 0, col_field = 0, row_stride = 0, col_stride = 0;
    int row_spacing = 0, col_spacing = 0, row_offset = 0, col_offset = 0;
    bool wrap = false;
    int total() const { return row_field * col_field; }
};

struct SubsetConfig {  // network/connection_config.cpp:112-134
    int from_row_start = 0, from_row_end = 0, from_col_start = 0,
        from_col_end = 0, to_row_start = 0, to_row_end = 0, to_col_start = 0,
        to_col_end = 0;
    int from_size() const {
        return (from_row_end - from_row_start) * (from_col_end - from_col_start);
    }
    int to_size() const {
        return (to_row_end - to_row_start) * (to_col_end - to_col_start);
    }
};

// One node of a layer's dendritic tree (network/dendritic_node.h:26-66,
// layer.cpp:135-152).  Internal nodes own an input register, numbered in
// the order the nodes are created (root 0); connections are leaves of the
// node they name and activate into its register.  A node's children run
// internal nodes first (created with the layer) and then its connections
// (created afterwards), as in the reference's child list.
struct Dendrite {
    std::string name;
    int opcode = SG_ADD;
    float init = 0.0f;
    int reg = 0;
    // second order (dendritic_node.h:53-59): the first connection is the
    // host, the later ones (its leaves) fold their synaptic activities into
    // a copy of the host's weights before the host activates
    bool second_order = false;
    Connection *host = nullptr;
    std::vector<Dendrite *> children;
    std::vector<Connection *> conns;
};

struct Layer {
    Structure *structure = nullptr;
    Props config;
    std::string name, neural_model;
    int rows = 0, cols = 0, size = 0;
    bool plastic = false;
    size_t id = 0;
    std::vector<Connection *> inputs, outputs;
    std::vector<std::unique_ptr<Dendrite>> dendrites;   // [0] is the root
    int registers = 1;                                  // input registers

    std::string param(const std::string &k, const std::string &def) const {
        return config.get(k, def);
    }
    Dendrite *dendrite(const std::string &name) const;
    int max_out_delay() const;
    std::string str() const;
};

struct Connection {
    Props config;
    std::string name;
    Layer *from = nullptr, *to = nullptr;
    int type = SG_FULLY_CONNECTED;
    int opcode = SG_ADD;
    int delay = 0;
    float max_weight = 1.0f;
    bool plastic = true, convolutional = false, sparse = false;
    bool recurrent = false;
    // randomized projection (connection_config.cpp:219): sources drawn at
    // random, the connection sparse (connection.cpp:138-139)
    bool randomized = false;
    int reg = 0;          // input register of its dendritic node
    std::string node = "root";                 // its dendritic node
    bool so_host = false, so_slave = false;    // second-order roles
    ArborConfig arbor;
    SubsetConfig subset;
    int num_weights = 0;
    // A sparse connection becomes SG_SPARSE with fewer weights when a state
    // is built (network/connection.cpp:219-222); the dense type and size
    // are kept so that a later state starts from them again.
    int dense_type = SG_FULLY_CONNECTED, dense_num_weights = 0;
    size_t id = 0;

    std::string param(const std::string &k, const std::string &def) const {
        return config.get(k, def);
    }
    std::string str() const;
    Props weight_config() const {
        Props *w = config.get_child("weight config");
        return w ? *w : Props();
    }
    // Host-layout matrix shape (connection.cpp:229-248).
    int matrix_rows() const;
    int matrix_columns() const;
};

struct Structure {
    std::string name;
    ClusterType cluster = CLUSTER_PARALLEL;
    std::vector<std::unique_ptr<Layer>> layers;
    std::vector<std::unique_ptr<Connection>> connections;
    Layer *layer(const std::string &n) const;
};

class Network {
  public:
    explicit Network(const Props &config);
    const Props &config() const { return config_; }
    std::vector<Layer *> layers() const;
    std::vector<Connection *> connections() const;
    Structure *structure(const std::string &name) const;
    Layer *layer(const std::string &name) const;
    Connection *connection(const std::string &name) const;
    const std::vector<std::unique_ptr<Structure>> &structures() const {
        return structures_;
    }
    int max_layer_size() const;

  private:
    void add_structure(const Props &sc);
    void add_connection(const Props &cc);
    Props config_;
    std::vector<std::unique_ptr<Structure>> structures_;
};

int connection_type_from(const std::string &s);
int opcode_from(const std::string &s);

}  // namespace sg

#endif
