// Error state, layout sizes and the cached device workspace.
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <tuple>

#include "common.hpp"

namespace sbh {

namespace {
thread_local char g_err[512] = {0};

struct WsEntry {
  void* ptr = nullptr;
  size_t bytes = 0;
};
std::mutex g_ws_mutex;
std::map<std::tuple<int, hipStream_t, int>, WsEntry> g_ws;
}  // namespace

int fail_invalid(const char* msg) {
  std::snprintf(g_err, sizeof(g_err), "%s", msg);
  return SBH_INVALID_ARGUMENT;
}

int fail_internal(hipError_t err) {
  std::snprintf(g_err, sizeof(g_err), "%s", hipGetErrorString(err));
  return SBH_INTERNAL;
}

int fail_internal_msg(const char* msg) {
  std::snprintf(g_err, sizeof(g_err), "%s", msg);
  return SBH_INTERNAL;
}

void clear_error() { g_err[0] = 0; }

void* workspace(hipStream_t stream, int slot, size_t bytes) {
  if (bytes == 0) bytes = 256;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return nullptr;
  std::lock_guard<std::mutex> lock(g_ws_mutex);
  WsEntry& e = g_ws[std::make_tuple(dev, stream, slot)];
  if (e.bytes >= bytes) return e.ptr;
  if (e.ptr != nullptr) {
    // Work queued on this stream may still read the old buffer.
    if (hipStreamSynchronize(stream) != hipSuccess) return nullptr;
    (void)hipFree(e.ptr);
    e.ptr = nullptr;
    e.bytes = 0;
  }
  size_t rounded = (bytes + (1u << 20) - 1) & ~static_cast<size_t>((1u << 20) - 1);
  void* p = nullptr;
  if (hipMalloc(&p, rounded) != hipSuccess) return nullptr;
  e.ptr = p;
  e.bytes = rounded;
  return p;
}

bool stream_capturing(hipStream_t stream) {
  hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
  if (hipStreamIsCapturing(stream, &st) != hipSuccess) return false;
  return st != hipStreamCaptureStatusNone;
}

int cached_workspace(hipStream_t stream, int slot, size_t bytes, char** out) {
  *out = nullptr;
  // A graph would keep the cached pointer, and a later eager call that needs
  // a larger buffer on the same stream handle frees it: replays would then
  // write freed memory.  Captured calls therefore bring their own workspace.
  if (stream_capturing(stream))
    return fail_invalid("a workspace buffer must be passed when the stream is being captured");
  *out = static_cast<char*>(workspace(stream, slot, bytes));
  if (*out == nullptr) return fail_internal_msg("workspace allocation failed");
  return SBH_OK;
}

}  // namespace sbh

extern "C" {

const char* sbh_last_error(void) { return sbh::g_err; }

size_t sbh_leaf3_size(void) { return sizeof(sbh::Leaf3); }
size_t sbh_node3_size(void) { return sizeof(sbh::Node3); }
size_t sbh_octree_level3_max_nodes(void) { return 1 + 8 * 8 * 8; }
size_t sbh_octree8_node_size(void) { return sizeof(sbh::Node8); }

}  // extern "C"
