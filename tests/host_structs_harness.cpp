// Host-only checks of the runtime's plan cache and the serving scratch pool,
// built with AddressSanitizer by tests/test_host_structs.py.  Exits 0 when
// every check passes; prints the failing check otherwise.
#include <atomic>
#include <cstdio>
#include <cstdlib>
#include <mutex>
#include <random>
#include <string>
#include <thread>
#include <vector>

#include "../eithot_amd/csrc/plan_cache.hpp"
#include "../eithot_amd/csrc/scratch_pool.hpp"

using eithot::LruCache;
using eithot::ScratchPool;

static int g_fail = 0;
#define CHECK(c)                                                     \
  do {                                                               \
    if (!(c)) {                                                      \
      std::printf("FAILED %s:%d: %s\n", __FILE__, __LINE__, #c);     \
      ++g_fail;                                                      \
    }                                                                \
  } while (0)

static std::atomic<int> g_live{0};

// stands in for a plan: owns host memory, counts live instances
struct FakePlan {
  std::vector<int> data;
  int tag;
  explicit FakePlan(int t) : data(1024, t), tag(t) { ++g_live; }
  ~FakePlan() { --g_live; }
};

static void lru_eviction_keeps_held_values() {
  LruCache<FakePlan> c(4);
  auto held = std::make_shared<const FakePlan>(7);
  c.put("held", held);
  std::weak_ptr<const FakePlan> watch = held;
  for (int i = 0; i < 20; ++i) c.put("k" + std::to_string(i), std::make_shared<const FakePlan>(i));
  CHECK(c.size() == 4);
  CHECK(c.get("held") == nullptr);  // evicted from the cache...
  CHECK(held->tag == 7);            // ...but alive for its holder
  for (int v : held->data) CHECK(v == 7);
  held.reset();
  CHECK(watch.expired());  // destroyed by its last holder
  CHECK(g_live == 4);
  auto out = c.clear();
  CHECK(c.size() == 0 && g_live == 4);
  out.clear();
  CHECK(g_live == 0);
}

static void lru_order_and_weight() {
  LruCache<FakePlan> c(3);
  for (int i = 0; i < 3; ++i) c.put("k" + std::to_string(i), std::make_shared<const FakePlan>(i));
  CHECK(c.get("k0") != nullptr);  // k0 becomes most recent; k1 is now the oldest
  c.put("k3", std::make_shared<const FakePlan>(3));
  CHECK(c.get("k1") == nullptr);
  CHECK(c.get("k0") != nullptr && c.get("k2") != nullptr && c.get("k3") != nullptr);
  // replacing a key keeps one entry
  c.put("k3", std::make_shared<const FakePlan>(30));
  CHECK(c.size() == 3 && c.get("k3")->tag == 30);
  c.clear();
  // weight cap: the newest entry stays even when it alone exceeds the cap
  LruCache<FakePlan> w(100, 1000);
  w.put("a", std::make_shared<const FakePlan>(1), 600);
  w.put("b", std::make_shared<const FakePlan>(2), 300);
  CHECK(w.size() == 2 && w.weight() == 900);
  w.put("c", std::make_shared<const FakePlan>(3), 300);  // a goes
  CHECK(w.get("a") == nullptr && w.size() == 2 && w.weight() == 600);
  w.put("d", std::make_shared<const FakePlan>(4), 5000);  // everything else goes
  CHECK(w.size() == 1 && w.get("d") != nullptr && w.weight() == 5000);
  w.clear();
  CHECK(g_live == 0);
}

// The runtime's pattern: threads look up or insert under a lock and use the
// value outside it while other threads overfill the cache.
static void lru_threads() {
  LruCache<FakePlan> c(8);
  std::mutex mu;
  std::atomic<int> bad{0};
  auto worker = [&](int seed) {
    std::mt19937 rng(seed);
    for (int it = 0; it < 4000; ++it) {
      const int k = (int)(rng() % 64);
      const std::string key = "k" + std::to_string(k);
      std::shared_ptr<const FakePlan> p;
      {
        std::lock_guard<std::mutex> lk(mu);
        p = c.get(key);
      }
      if (!p) {
        auto q = std::make_shared<const FakePlan>(k);
        std::vector<std::shared_ptr<const FakePlan>> ev;
        std::lock_guard<std::mutex> lk(mu);
        p = c.get(key);
        if (!p) {
          p = q;
          ev = c.put(key, q);
        }
      }
      long sum = 0;
      for (int v : p->data) sum += v;  // use outside the lock
      if (p->tag != k || sum != 1024L * k) ++bad;
    }
  };
  std::vector<std::thread> ts;
  for (int t = 0; t < 6; ++t) ts.emplace_back(worker, t + 1);
  for (auto& t : ts) t.join();
  CHECK(bad == 0);
  c.clear();
  CHECK(g_live == 0);
}

static void pool_cap_and_granules() {
  ScratchPool p;
  p.init(1000);  // 3 whole granules: 768 usable bytes
  CHECK(p.cap == 768);
  CHECK(!p.fits(900));  // refused at submit: could never be allocated
  CHECK(!p.fits(769));
  CHECK(p.fits(768) && p.fits(700) && p.fits(0));
  const int64_t a = p.alloc(700);
  CHECK(a == 0);
  CHECK(p.alloc(1) == -1);  // full
  p.release(a, 700);
  CHECK(p.free_bytes() == 768);
  const int64_t x = p.alloc(10), y = p.alloc(300), z = p.alloc(200);
  CHECK(x == 0 && y == 256 && z == -1);  // y took granules 1 and 2
  p.release(x, 10);
  CHECK(p.alloc(256) == 0);
  p.release(0, 256);
  p.release(y, 300);
  CHECK(p.free_.size() == 1 && p.free_bytes() == 768);  // coalesced
  ScratchPool e;
  e.init(100);  // under one granule: only zero-scratch requests fit
  CHECK(e.cap == 0 && e.fits(0) && !e.fits(1));
}

// random allocations against a byte map: no overlap, full coalescing
static void pool_random() {
  ScratchPool p;
  const int64_t cap = 256 * 97 + 55;
  p.init(cap);
  std::vector<int> owner(p.cap, -1);
  std::vector<std::pair<int64_t, int64_t>> live;
  std::mt19937 rng(5);
  for (int it = 0; it < 20000; ++it) {
    if (live.empty() || rng() % 2) {
      const int64_t n = 1 + rng() % 4000;
      const int64_t off = p.alloc(n);
      if (off < 0) continue;
      const int64_t g = ScratchPool::granule(n);
      CHECK(off % 256 == 0 && off + g <= p.cap);
      for (int64_t b = off; b < off + g && b < p.cap; ++b) {
        CHECK(owner[b] == -1);
        owner[b] = it;
      }
      live.push_back({off, n});
    } else {
      const size_t i = rng() % live.size();
      const auto [off, n] = live[i];
      for (int64_t b = off; b < off + ScratchPool::granule(n); ++b) owner[b] = -1;
      p.release(off, n);
      live[i] = live.back();
      live.pop_back();
    }
  }
  for (const auto& [off, n] : live) p.release(off, n);
  CHECK(p.free_.size() == 1 && p.free_bytes() == p.cap);
}

int main() {
  lru_eviction_keeps_held_values();
  lru_order_and_weight();
  lru_threads();
  pool_cap_and_granules();
  pool_random();
  if (g_fail) return 1;
  std::printf("host structs ok\n");
  return 0;
}
