// Optional phase timing of the adaptive-tree pipelines: HIP events recorded
// on the call's stream around the tree build(s) and around the walk kernel,
// read back by sbh_profile_collect().  Kinds: 0 gravity forward (single set
// or batch), 1 gravity adjoint, 2 winding numbers, 3 stochastic walks.
#pragma once

#include <hip/hip_runtime.h>

namespace sbh {
namespace prof {

constexpr int kKinds = 4;

struct Rec {
  hipEvent_t ev[4];  // build start, build end, walk kernel start, walk kernel end
  int kind;
};

// Creates the events and records the first one when profiling is enabled;
// false (and no event left alive) otherwise.
bool begin(hipStream_t s, Rec* r, int kind);
// Queues a completed record for sbh_profile_collect (which destroys it).
void end(const Rec& r);
// Destroys the events of a record that will not be queued.
void destroy(Rec& r);

// Owns one record's events from begin() until commit(); a call that returns
// early (an error, an empty input) drops them in the destructor.
class Scope {
 public:
  Scope(hipStream_t s, int kind) { on_ = begin(s, &r_, kind); }
  ~Scope() {
    if (on_) destroy(r_);
  }
  Scope(const Scope&) = delete;
  Scope& operator=(const Scope&) = delete;
  bool on() const { return on_; }
  Rec& rec() { return r_; }
  void commit() {
    if (on_) end(r_);
    on_ = false;
  }

 private:
  Rec r_{};
  bool on_ = false;
};

}  // namespace prof
}  // namespace sbh
