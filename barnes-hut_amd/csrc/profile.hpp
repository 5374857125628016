// Optional phase timing of the adaptive-tree pipelines: HIP events recorded
// on the call's stream around the tree build(s) and around the walk kernel,
// read back by sbh_profile_collect().  Kinds: 0 gravity forward (single set
// or batch), 1 gravity adjoint, 2 winding numbers.
#pragma once

#include <hip/hip_runtime.h>

namespace sbh {
namespace prof {

constexpr int kKinds = 3;

struct Rec {
  hipEvent_t ev[4];  // build start, build end, walk kernel start, walk kernel end
  int kind;
};

// Creates the events and records the first one when profiling is enabled;
// false (and nothing recorded) otherwise.
bool begin(hipStream_t s, Rec* r, int kind);
// Queues a completed record for sbh_profile_collect.
void end(const Rec& r);

}  // namespace prof
}  // namespace sbh
