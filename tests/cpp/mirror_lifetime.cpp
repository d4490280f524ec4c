// Stress test of the single-point mirror's lifetime (mirror_cache.hpp) on
// the CPU: evaluator threads hold the current mirror while a setter thread
// keeps marking the spline dirty, as concurrent set_* calls do.  A mirror
// must stay alive while any call holds it, every replaced mirror must be
// freed exactly once, and only the current one may be left.  Built with
// AddressSanitizer, so a use after free aborts.
#include <assert.h>
#include <pthread.h>
#include <stdio.h>

#include <atomic>
#include <thread>
#include <vector>

#include "mirror_cache.hpp"

struct Fake {
  unsigned magic = 0x4d495252u;
  long serial;
};

static std::atomic<long> built{0}, freed{0};

int main() {
  pthread_mutex_t lock;
  pthread_mutex_init(&lock, nullptr);
  int dirty = 1;
  void *slot = nullptr;
  auto build = [&]() -> void * {
    Fake *f = new Fake;
    f->serial = built++;
    return f;
  };
  auto destroy = [&](void *p) {
    Fake *f = (Fake *)p;
    assert(f->magic == 0x4d495252u);
    f->magic = 0;
    delete f;
    freed++;
  };
  std::atomic<bool> stop{false};
  std::atomic<long> calls{0};
  std::vector<std::thread> th;
  for (int w = 0; w < 6; ++w)
    th.emplace_back([&] {
      while (!stop) {
        esk::MirrorRef *r = esk::mirror_acquire(&lock, &dirty, &slot, build, destroy);
        Fake *f = (Fake *)r->table;
        for (int k = 0; k < 50; ++k) assert(f->magic == 0x4d495252u);   // the call's evaluation
        std::this_thread::yield();
        assert(f->magic == 0x4d495252u);
        esk::mirror_release(&lock, r, destroy);
        calls++;
      }
    });
  th.emplace_back([&] {
    for (int i = 0; i < 20000; ++i) {
      pthread_mutex_lock(&lock);
      dirty = 1;
      pthread_mutex_unlock(&lock);
      if (i % 64 == 0) std::this_thread::yield();
    }
    stop = true;
  });
  for (auto &t : th) t.join();
  // exactly the current mirror is left, with no users
  esk::MirrorRef *cur = (esk::MirrorRef *)slot;
  assert(cur && cur->users == 0 && !cur->retired);
  assert(built - freed == 1);
  esk::mirror_free(slot, destroy);
  assert(built == freed);
  printf("mirror lifetime ok: %ld calls, %ld mirrors built and freed\n", calls.load(), built.load());
  return 0;
}
