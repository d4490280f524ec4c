// mirror_cache.hpp -- lifetime of the device mirror that the single-point
// API evaluates (host code only; no HIP).
//
// A host spline keeps one current mirror of its coefficients.  set_* (or an
// invalidate) marks the spline dirty under its lock; the next call builds a
// new mirror and makes it current.  A call holds a reference to the mirror
// it evaluates, so a mirror replaced while other calls still use it is only
// retired, and the last of those calls frees it.
#pragma once
#include <pthread.h>

namespace esk {

struct MirrorRef {
  void *table = nullptr;   // the device mirror
  int users = 0;           // calls holding this mirror
  bool retired = false;    // replaced: free when users drops to 0
};

// Reference to the current mirror behind *slot, built with build() (which
// returns nullptr on failure) when *dirty or there is none.  Returns nullptr
// if the build fails (nothing changes then).  destroy(table) frees a mirror.
template <class Build, class Destroy>
MirrorRef *mirror_acquire(pthread_mutex_t *lock, int *dirty, void **slot, Build build, Destroy destroy) {
  MirrorRef *drop = nullptr;
  pthread_mutex_lock(lock);
  MirrorRef *cur = (MirrorRef *)*slot;
  if (*dirty || !cur) {
    void *m = build();
    if (!m) {
      pthread_mutex_unlock(lock);
      return nullptr;
    }
    MirrorRef *fresh = new MirrorRef;
    fresh->table = m;
    if (cur) {
      if (cur->users == 0) drop = cur;
      else cur->retired = true;
    }
    *slot = fresh;
    *dirty = 0;
    cur = fresh;
  }
  cur->users++;
  pthread_mutex_unlock(lock);
  if (drop) {
    destroy(drop->table);
    delete drop;
  }
  return cur;
}

template <class Destroy>
void mirror_release(pthread_mutex_t *lock, MirrorRef *r, Destroy destroy) {
  pthread_mutex_lock(lock);
  const bool last = --r->users == 0 && r->retired;
  pthread_mutex_unlock(lock);
  if (last) {
    destroy(r->table);
    delete r;
  }
}

// The spline is being destroyed: no call may still hold its mirror.
template <class Destroy>
void mirror_free(void *slot_value, Destroy destroy) {
  MirrorRef *r = (MirrorRef *)slot_value;
  if (!r) return;
  destroy(r->table);
  delete r;
}

}  // namespace esk
