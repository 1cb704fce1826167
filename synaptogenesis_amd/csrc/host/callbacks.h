// Named callback registry (util/callback_manager.{h,cpp}): the C-ABI's
// add_*_callback entry points store raw function addresses by name, and
// weight configs / callback modules look them up.
#ifndef SG_CALLBACKS_H
#define SG_CALLBACKS_H

#include <map>
#include <mutex>
#include <string>

#include "props.h"

namespace sg {

typedef void (*IoCallback)(int, int, void *);
typedef void (*WeightCallback)(int, int, void *);
typedef void (*IndicesCallback)(int, int, void *, void *, void *, void *);
typedef void (*DistanceCallback)(int, int, void *, void *);
typedef void (*DelayCallback)(int, int, void *, void *);

class Callbacks {
  public:
    static Callbacks &get() {
        static Callbacks c;
        return c;
    }
    template <class F>
    void add(std::map<std::string, F> &m, const std::string &n, F f) {
        std::lock_guard<std::mutex> g(mu_);
        m[n] = f;
    }
    template <class F>
    F find(const std::map<std::string, F> &m, const std::string &n,
           const char *kind) {
        std::lock_guard<std::mutex> g(mu_);
        auto it = m.find(n);
        if (it == m.end())
            throw Error(std::string("Unrecognized ") + kind + " callback: " +
                        n);
        return it->second;
    }
    std::map<std::string, IoCallback> io;
    std::map<std::string, WeightCallback> weight;
    std::map<std::string, IndicesCallback> indices;
    std::map<std::string, DistanceCallback> distance;
    std::map<std::string, DelayCallback> delay;

  private:
    std::mutex mu_;
};

}  // namespace sg

#endif
