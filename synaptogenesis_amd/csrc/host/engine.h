// Step scheduler: runs a State for a number of iterations on one GPU stream
// in the order of the reference's single-thread loop and parallel cluster
// (engine/engine.cpp:226-315, engine/cluster/parallel_cluster.cpp:27-131):
//
//   1. layers without environment input: clear (or init) the input, then
//      every input connection's activator in connection order;
//   2. environment input: callback modules fill host buffers, which are
//      copied into the input of input layers;
//   3. input layers: init (added) and activators;
//   4. every layer's neuron kernel;
//   5. if learning, every plastic connection's updater;
//   6. output: device -> host copies and callback modules.
//
// The updater of a plastic Izhikevich conductance connection is deferred
// into the next step's activation pass of the same connection (one fused
// read/write of its per-synapse arrays), also across runs: the next run
// starts with the fused pass, and any read of the state from the host
// (getters, save) first applies what is still deferred, so the state a
// caller sees is the reference's.
//
// Streams: each structure's work runs on a stream of its own (and the
// activators that write a layer's input on a second stream of it); a
// scheduler orders launches on different streams by the arrays they read
// and write, with events only where one stream's launch depends on
// another's.  Runs without per-step host I/O capture their steady-state
// steps into HIP graphs (kept with the state for later runs).
#ifndef SG_ENGINE_H
#define SG_ENGINE_H

#include <atomic>
#include <memory>
#include <string>
#include <vector>

#include "props.h"
#include "state.h"

namespace sg {

class Engine {
  public:
    Engine(Network *net, const Props *environment, State *state)
        : net_(net), env_(environment), state_(state) {}
    // Runs and returns a report (caller owns it).  Throws on error.
    Props *run(const Props &args);

    static void interrupt();

  private:
    Network *net_;
    const Props *env_;
    State *state_;
};

// Applies every update still deferred into a next activation (a plastic
// conductance connection's update of the last step run), so that the
// state's arrays are the reference's after that step.  State::to_host
// calls it before any host read.
void flush_pending_updates(State &state);

}  // namespace sg

#endif
