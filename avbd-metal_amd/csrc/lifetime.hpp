// lifetime.hpp -- process-exit safety for the library's handles.
//
// Every solver, network and snapshot is registered while it is alive.  The
// first registration installs an atexit handler; it is installed after the
// HIP runtime has initialised (the create call has made HIP calls by then),
// so it runs before the runtime's own exit handlers and releases every
// handle still alive while the runtime can still take the calls.  A destroy
// call arriving after that (a garbage collector finalising its objects late)
// finds its handle unregistered and returns without touching the GPU.
//
// The registry is heap-allocated and never freed, so no static destructor
// can run before a late destroy call.
#pragma once
#include <stdlib.h>

#include <mutex>
#include <vector>

namespace avbd {

enum class HandleKind : int { Snapshot = 0, Solver = 1, Network = 2 };

struct LiveHandle {
  void* p;
  HandleKind kind;
};

struct LiveRegistry {
  std::mutex mu;
  std::vector<LiveHandle> live;
  bool hooked = false;
  bool exiting = false;
};

inline LiveRegistry& live_registry() {
  static LiveRegistry* r = new LiveRegistry();
  return *r;
}

// releases every handle still alive at exit (defined in solver.hip:
// snapshots, then solvers, then networks)
void release_all_at_exit();

inline void live_add(void* p, HandleKind k) {
  LiveRegistry& r = live_registry();
  std::lock_guard<std::mutex> g(r.mu);
  r.live.push_back({p, k});
  if (!r.hooked) {
    r.hooked = true;
    atexit(release_all_at_exit);
  }
}

// true when p was alive (and is now unregistered): the caller may release it
inline bool live_remove(void* p) {
  LiveRegistry& r = live_registry();
  std::lock_guard<std::mutex> g(r.mu);
  for (size_t i = 0; i < r.live.size(); i++)
    if (r.live[i].p == p) {
      r.live.erase(r.live.begin() + (long)i);
      return true;
    }
  return false;
}

inline bool live_contains(const void* p) {
  LiveRegistry& r = live_registry();
  std::lock_guard<std::mutex> g(r.mu);
  for (const LiveHandle& h : r.live)
    if (h.p == p) return true;
  return false;
}

}  // namespace avbd
