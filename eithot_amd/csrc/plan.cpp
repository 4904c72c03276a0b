// Host-side planner: decomposer, stage lowering, estimators.
#include "plan.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <new>
#include <sstream>
#include <thread>
#include <tuple>

#include "eithot_status.h"

namespace eithot {

int64_t gcd64(int64_t a, int64_t b) {
  while (b) {
    int64_t t = a % b;
    a = b;
    b = t;
  }
  return a;
}

int64_t modinv64(int64_t a, int64_t m) {
  if (m <= 1) return 0;
  // extended Euclid on (a mod m, m)
  __int128 t = 0, nt = 1, r = m, nr = a % m;
  while (nr != 0) {
    __int128 q = r / nr;
    __int128 tmp = t - q * nt;
    t = nt;
    nt = tmp;
    tmp = r - q * nr;
    r = nr;
    nr = tmp;
  }
  if (t < 0) t += m;
  return (int64_t)t;
}

// ============================================================================
// Decomposer.
//
// The reference generates each round with permutation_util.cpp:300-374 and
// picks pre- or post-order with Nd_transpose.cpp:146-157.  Written here as
// its equivalent block form: a round takes the current dim order `cur`, finds
// the anchor mismatch against the target, and swaps two adjacent blocks X|Y
// -> Y|X; the untouched prefix P and suffix S are merged into single dims.
//   pre-order  (permutation_util.cpp:146-180): u = first mismatch,
//     X = cur[u, t), Y = the longest run cur[t, t+L) equal to target[u, u+L)
//     where t is the position of target[u] in cur.
//   post-order (permutation_util.cpp:183-230): u = last mismatch,
//     Y = cur[t+1, u], X = the longest run cur[t-L+1, t] equal to
//     target[u-L+1, u].
// ============================================================================
namespace {

struct Round {
  bool done = false;
  Stage st;
  std::vector<int> next;
};

Round make_round(bool post, const std::vector<int>& cur, const std::vector<int>& target,
                 const int64_t* dims) {
  const int r = (int)cur.size();
  Round R;
  int u = -1;
  if (!post) {
    for (int i = 0; i < r; ++i)
      if (cur[i] != target[i]) { u = i; break; }
  } else {
    for (int i = r - 1; i >= 0; --i)
      if (cur[i] != target[i]) { u = i; break; }
  }
  if (u < 0 || (post && u == 0)) {
    R.done = true;
    return R;
  }
  const int t = (int)(std::find(cur.begin(), cur.end(), target[u]) - cur.begin());
  int x0, y0, y1;  // P = [0,x0) X = [x0,y0) Y = [y0,y1) S = [y1,r)
  if (!post) {
    int L = 1;
    while (u + L < r && t + L < r && target[u + L] == cur[t + L]) ++L;
    x0 = u;
    y0 = t;
    y1 = t + L;
  } else {
    int L = 1;
    while (u - L >= 0 && t - L >= 0 && target[u - L] == cur[t - L]) ++L;
    x0 = t - L + 1;
    y0 = t + 1;
    y1 = u + 1;
  }
  auto prod = [&](int lo, int hi) {
    int64_t p = 1;
    for (int i = lo; i < hi; ++i) p *= dims[cur[i] - 1];
    return p;
  };
  Stage& s = R.st;
  s.P = prod(0, x0);
  s.X = prod(x0, y0);
  s.Y = prod(y0, y1);
  s.S = prod(y1, r);
  const bool hasP = x0 > 0, hasS = y1 < r;
  int n = 0;
  if (hasP) s.dims[n++] = s.P;
  s.dims[n++] = s.X;
  s.dims[n++] = s.Y;
  if (hasS) s.dims[n++] = s.S;
  s.rank = n;
  // output order P, Y, X, S expressed as 1-based positions of the input blocks
  int xi = hasP ? 2 : 1, yi = xi + 1;
  n = 0;
  if (hasP) s.perm[n++] = 1;
  s.perm[n++] = yi;
  s.perm[n++] = xi;
  if (hasS) s.perm[n++] = yi + 1;
  R.next.reserve(r);
  for (int i = 0; i < x0; ++i) R.next.push_back(cur[i]);
  for (int i = y0; i < y1; ++i) R.next.push_back(cur[i]);
  for (int i = x0; i < y0; ++i) R.next.push_back(cur[i]);
  for (int i = y1; i < r; ++i) R.next.push_back(cur[i]);
  return R;
}

// Nd_transpose.cpp:104-113: 1 when a stage cannot run as a single
// linearized pass under the reference's thresholds.
int not_linearizable(const Stage& s) {
  if (s.rank == 3) {
    if (s.perm[2] == 3 && s.dims[2] >= 8) return 0;
    if (s.perm[0] == 1 && s.dims[0] >= 48) return 0;
  } else if (s.rank == 4) {
    if (s.perm[3] == 4 && s.dims[3] >= 4) return 0;
    if (s.perm[0] == 1 && s.dims[0] >= 96) return 0;
  }
  return 1;
}

int count_not_linearizable(bool post, const std::vector<int>& target, const int64_t* dims) {
  const int r = (int)target.size();
  std::vector<int> cur(r);
  for (int i = 0; i < r; ++i) cur[i] = i + 1;
  int cnt = 0;
  for (int round = 0; round < r; ++round) {
    Round R = make_round(post, cur, target, dims);
    if (R.done) break;
    cnt += not_linearizable(R.st);
    cur = R.next;
    if (cur == target) break;
  }
  return cnt;
}

}  // namespace

bool decompose(const int64_t* dims, const int* perm1, int rank, int reorder,
               std::vector<Stage>& out) {
  out.clear();
  if (rank < 1 || rank > 64) return false;
  std::vector<int> target(perm1, perm1 + rank), seen(rank + 1, 0);
  for (int v : target) {
    if (v < 1 || v > rank || seen[v]) return false;
    seen[v] = 1;
  }
  bool post;
  if (reorder == REORDER_PRE) post = false;
  else if (reorder == REORDER_POST) post = true;
  else post = !(count_not_linearizable(false, target, dims) < count_not_linearizable(true, target, dims));
  std::vector<int> cur(rank);
  for (int i = 0; i < rank; ++i) cur[i] = i + 1;
  for (int round = 0; round < rank; ++round) {
    // the reference stops as soon as either generator reports "matched"
    Round pre = make_round(false, cur, target, dims);
    Round pst = make_round(true, cur, target, dims);
    if (pre.done || pst.done) break;
    Round& R = post ? pst : pre;
    out.push_back(R.st);
    cur = R.next;
    if (cur == target) break;
  }
  return true;
}

// ============================================================================
// Lowering of one batched 2-D transpose (B batches of R x C units of U bytes,
// C fastest) to device passes.
// ============================================================================

int64_t Op::footprint_units() const {
  switch (kind) {
    case OP_TILE: return R * C;
    case OP_COL: return R * W;
    case OP_ROW: return C;
    case OP_CYCLE: return 1;  // scratch: one saved cell per boundary slot
    case OP_BLOCK: return R;
    case OP_SEGTILE: return R * C;
  }
  return 0;
}

int64_t Op::num_footprints() const {
  const int64_t slices = U / us;
  switch (kind) {
    case OP_TILE: return B * slices;
    case OP_COL: return B * ((C + W - 1) / W) * slices;
    case OP_ROW: return B * R * slices;
    case OP_CYCLE: return cyc ? B * (int64_t)cyc->slots.size() : 0;
    case OP_BLOCK: return B;
    case OP_SEGTILE: return B;
  }
  return 0;
}

namespace {
// x / d for x, d < 2^32 by one multiply: with M = floor((2^64 - 1) / d) + 1,
// floor(x * M / 2^64) = floor(x / d) because x * d < 2^64.
struct HostDiv {
  uint64_t d = 1, M = 0;
  explicit HostDiv(uint64_t dv) : d(dv ? dv : 1), M(d > 1 ? ~0ull / d + 1 : 0) {}
  uint64_t div(uint64_t x) const {
    return d == 1 ? x : (uint64_t)(((unsigned __int128)x * M) >> 64);
  }
};

struct MapDiv {
  HostDiv bc, m, R;
  uint64_t C;
};

std::mutex g_cyc_mu;
std::map<std::vector<int64_t>, std::shared_ptr<CycleTable>> g_cyc_cache;  // recent tables

// Tables over at least this many cells are walked on several threads.
constexpr uint64_t kParallelWalkCells = 1ull << 20;

// fn(w) for w in [0, n) on up to n threads; a share whose thread cannot be
// started runs on the calling thread.
template <typename Fn>
void run_shares(unsigned n, const Fn& fn) {
  std::vector<std::thread> th;
  std::vector<unsigned> inline_shares;
  for (unsigned w = 1; w < n; ++w) {
    try {
      th.emplace_back(fn, w);
    } catch (...) {
      inline_shares.push_back(w);
    }
  }
  fn(0u);
  for (unsigned w : inline_shares) fn(w);
  for (std::thread& x : th) x.join();
}

// The cycles that hold a hashed boundary position (about one position in
// 64) are walked on several threads, piece by piece: a piece starts at a
// boundary and follows p -> src(p) up to the next boundary.  Every position
// of such a cycle lies on exactly one piece.  One thread then chains the
// pieces of each cycle and cuts the cycle into segments of lmax positions,
// exactly as a single walk would (a segment start inside a piece is found
// by stepping from the piece's start).  Cycles without a boundary stay
// unmarked in `seen` for the sequential walk.
template <typename Src>
void walk_parallel(const Src& src, uint64_t N, int64_t lmax, std::vector<uint64_t>& seen,
                   CycleTable& t) {
  constexpr uint64_t kPiece = 64;
  auto boundary = [&](uint64_t p) { return ((p * 0x9E3779B97F4A7C15ull) >> 40) % kPiece == 0; };
  struct Piece {
    uint32_t start, len, next;  // next: the start of the following piece
  };
  const unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  std::vector<std::vector<Piece>> out(nt);
  auto work = [&](unsigned w) {
    const uint64_t lo = N * w / nt, hi = N * (w + 1) / nt;
    std::vector<Piece>& v = out[w];
    for (uint64_t p = lo; p < hi; ++p) {
      if (!boundary(p)) continue;
      uint64_t q = p, k = 0;
      do {
        __atomic_fetch_or(&seen[q >> 6], 1ull << (q & 63), __ATOMIC_RELAXED);
        q = src(q);
        ++k;
      } while (q != p && !boundary(q));
      v.push_back({(uint32_t)p, (uint32_t)k, (uint32_t)q});
    }
  };
  run_shares(nt, work);
  std::vector<Piece> all;
  for (const auto& v : out) all.insert(all.end(), v.begin(), v.end());
  std::sort(all.begin(), all.end(), [](const Piece& a, const Piece& b) { return a.start < b.start; });
  // index of each piece's successor, on the threads
  std::vector<uint32_t> nxt(all.size());
  auto link = [&](unsigned w) {
    const size_t lo = all.size() * w / nt, hi = all.size() * (w + 1) / nt;
    for (size_t i = lo; i < hi; ++i)
      nxt[i] = (uint32_t)(std::lower_bound(all.begin(), all.end(), all[i].next,
                                           [](const Piece& a, uint32_t x) { return a.start < x; }) -
                          all.begin());
  };
  run_shares(nt, link);
  const uint64_t L = (uint64_t)lmax;
  std::vector<uint8_t> used(all.size(), 0);
  std::vector<uint32_t> chain;
  std::vector<uint64_t> at_off;  // cycle position of each chained piece's start
  std::vector<uint32_t> starts;
  for (size_t i0 = 0; i0 < all.size(); ++i0) {
    if (used[i0]) continue;
    chain.clear();
    at_off.clear();
    uint64_t total = 0;
    for (uint32_t i = (uint32_t)i0; !used[i]; i = nxt[i]) {
      used[i] = 1;
      chain.push_back(i);
      at_off.push_back(total);
      total += all[i].len;
    }
    if (total == 1) continue;  // a fixed point
    if (total <= L) {
      // a whole cycle records its last position (the end of its last piece)
      const Piece& lp = all[chain.back()];
      uint64_t last = lp.start;
      for (uint32_t x = 1; x < lp.len; ++x) last = src(last);
      t.segs.push_back({all[i0].start, (uint32_t)total, -1, (uint32_t)last});
      continue;
    }
    // segment starts every lmax positions along the cycle from all[i0].start,
    // each stepped to from the start of the piece that holds it (threads)
    const uint64_t n64 = (total + L - 1) / L;
    starts.assign(n64, 0);
    auto cut = [&](unsigned w, unsigned parts) {
      const uint64_t lo = n64 * w / parts, hi = n64 * (w + 1) / parts;
      for (uint64_t c = lo; c < hi; ++c) {
        const uint64_t at = c * L;
        const size_t k = (size_t)(std::upper_bound(at_off.begin(), at_off.end(), at) -
                                  at_off.begin()) - 1;
        uint64_t q = all[chain[k]].start;
        for (uint64_t x = at_off[k]; x < at; ++x) q = src(q);
        starts[c] = (uint32_t)q;
      }
    };
    const unsigned parts = n64 >= 4096 ? nt : 1;
    run_shares(parts, [&](unsigned w) { cut(w, parts); });
    const uint32_t n = (uint32_t)starts.size(), first_slot = (uint32_t)t.slots.size();
    for (uint32_t k = 0; k < n; ++k) {
      const uint64_t len = k + 1 < n ? L : total - (uint64_t)(n - 1) * L;
      t.slots.push_back(starts[k]);
      t.segs.push_back({starts[k], (uint32_t)len, (int32_t)(first_slot + (k + 1) % n), 0});
    }
  }
}
}  // namespace

std::shared_ptr<CycleTable> build_cycle_table_maps(const std::vector<CycleMap>& maps, int64_t N,
                                                   int64_t lmax) {
  std::vector<int64_t> key = {N, lmax};
  for (const CycleMap& M : maps) key.insert(key.end(), {M.bc, M.R, M.C, M.m});
  {
    std::lock_guard<std::mutex> lk(g_cyc_mu);
    auto it = g_cyc_cache.find(key);
    if (it != g_cyc_cache.end()) return it->second;
  }
  auto t = std::make_shared<CycleTable>();
  t->maps = maps;
  t->lmax = lmax;
  if (N <= 0 || (uint64_t)N >= (1ull << 32) || lmax < 1 || maps.empty()) return t;
  std::vector<MapDiv> md;
  for (const CycleMap& M : maps)
    md.push_back(MapDiv{HostDiv((uint64_t)M.bc), HostDiv((uint64_t)M.m), HostDiv((uint64_t)M.R),
                        (uint64_t)M.C});
  auto src = [&](uint64_t p) {
    for (size_t k = md.size(); k-- > 0;) {
      const MapDiv& M = md[k];
      const uint64_t b = M.bc.div(p), q = p - b * M.bc.d;
      const uint64_t c = M.m.div(q), x = q - c * M.m.d;
      const uint64_t j = M.R.div(c), i = c - j * M.R.d;
      p = b * M.bc.d + (i * M.C + j) * M.m.d + x;
    }
    return p;
  };
  std::vector<uint64_t> seen(((size_t)N + 63) / 64, 0);
  if ((uint64_t)N >= kParallelWalkCells) walk_parallel(src, (uint64_t)N, lmax, seen, *t);
  // one walk per remaining cycle over the visited bitset; a cycle longer
  // than lmax is cut into segments of lmax positions (the last one
  // shorter), each starting at a recorded position with a boundary slot
  std::vector<uint32_t> starts;
  for (uint64_t p0 = 0; p0 < (uint64_t)N; ++p0) {
    if (seen[p0 >> 6] >> (p0 & 63) & 1) continue;
    starts.clear();
    uint64_t k = 0, q = p0, left = 0, last = p0;
    do {
      if (left == 0) {
        starts.push_back((uint32_t)q);
        left = (uint64_t)lmax;
      }
      --left;
      seen[q >> 6] |= 1ull << (q & 63);
      last = q;
      q = src(q);
      ++k;
    } while (q != p0);
    if (k == 1) continue;
    if (starts.size() == 1) {
      t->segs.push_back({(uint32_t)p0, (uint32_t)k, -1, (uint32_t)last});
      continue;
    }
    const uint32_t n = (uint32_t)starts.size(), first_slot = (uint32_t)t->slots.size();
    for (uint32_t i = 0; i < n; ++i) {
      const uint64_t len = i + 1 < n ? (uint64_t)lmax : k - (uint64_t)(n - 1) * (uint64_t)lmax;
      t->slots.push_back(starts[i]);
      t->segs.push_back({starts[i], (uint32_t)len, (int32_t)(first_slot + (i + 1) % n), 0});
    }
  }
  std::lock_guard<std::mutex> lk(g_cyc_mu);
  if (g_cyc_cache.size() >= 16) g_cyc_cache.clear();
  g_cyc_cache[key] = t;
  return t;
}

std::shared_ptr<CycleTable> build_cycle_table(int64_t R, int64_t C, int64_t lmax) {
  return build_cycle_table_maps({CycleMap{R * C, R, C, 1}}, R * C, lmax);
}

namespace {

int64_t pow2_factor(int64_t x, int64_t cap) {
  int64_t w = 1;
  while (w < cap && x % (w * 2) == 0) w *= 2;
  return w;
}

const std::vector<int64_t>& divisors(int64_t x);  // ascending, memoized

// Largest slice of a U-byte unit such that `units` slices fit `budget`
// bytes, keeping the unit's power-of-two alignment (so the word width of the
// copy does not shrink) and at least `min_seg` contiguous bytes.  0 if none.
int64_t pick_slice(int64_t units, int64_t U, int64_t budget, int64_t min_seg) {
  if (units <= 0) return 0;
  if (units * U <= budget) return U;
  const int64_t align = pow2_factor(U, 16);
  const int64_t cap = budget / units;
  if (cap < min_seg) return 0;
  // the largest divisor of U that is a multiple of `align`, at most cap
  const std::vector<int64_t>& dv = divisors(U);
  for (auto it = dv.rbegin(); it != dv.rend(); ++it)
    if (*it % align == 0 && *it <= cap && *it >= min_seg) return *it;
  return 0;
}

// Tile footprint (R rows of C slices, padded rows): whole units if they fit
// the footprint target; else the largest aligned slice of at least slice_min
// bytes that does; else whole units if they fit LDS; else the largest
// aligned slice of at least min_segment bytes that fits LDS.  0 if none.
int64_t pick_tile_slice(int64_t R, int64_t C, int64_t U, const DeviceProfile& prof) {
  auto fp = [&](int64_t us) { return R * tile_row_stride_bound(C * us, us); };
  const int64_t align = pow2_factor(U, 16);
  auto largest = [&](int64_t cap, int64_t min_seg) {
    const std::vector<int64_t>& dv = divisors(U);
    for (auto it = dv.rbegin(); it != dv.rend() && *it >= min_seg; ++it)
      if (*it % align == 0 && fp(*it) <= cap) return *it;
    return (int64_t)0;
  };
  const int64_t target = prof.tile_fp_target > 0 ? std::min(prof.tile_fp_target, prof.lds_max)
                                                 : prof.lds_max;
  if (fp(U) <= target) return U;
  if (int64_t us = largest(target, std::max(prof.slice_min, prof.min_segment))) return us;
  if (fp(U) <= prof.lds_max) return U;
  return largest(prof.lds_max, prof.min_segment);
}

// Slice of a staged TILE: the largest aligned slice (at least min(U, 256)
// bytes) whose footprint fits both the scratch budget and the staging chunk.
int64_t pick_staged_tile_slice(int64_t R, int64_t C, int64_t U, int64_t budget,
                               const DeviceProfile& prof) {
  if (prof.stage_chunk > 0) {
    const int64_t min_seg = std::min<int64_t>(U, std::max(prof.slice_min, prof.min_segment));
    if (int64_t us = pick_slice(R * C, U, std::min(budget, prof.stage_chunk), min_seg)) return us;
  }
  return pick_slice(R * C, U, budget, std::min<int64_t>(U, 256));
}


// Largest divisor of U whose `units` slices fit `cap` bytes, whatever its
// alignment (narrow slices cost word width, but always exist).  0 if none.
int64_t any_slice(int64_t units, int64_t U, int64_t cap) {
  const std::vector<int64_t>& dv = divisors(U);
  for (auto it = dv.rbegin(); it != dv.rend(); ++it)
    if (units * *it <= cap) return *it;
  return 0;
}

// A Catanzaro column or row pass moving `us`-byte cells: word-granular
// gathers of small cells run far below HBM speed (0.4-1 TB/s measured for
// 2- and 4-byte units), staged passes cross HBM twice.
double cat_pass_cost(int64_t us, bool staged) {
  const double f = us >= 16 ? 1.0 : us >= 8 ? 2.0 : us >= 4 ? 4.0 : 7.0;
  return staged ? 2.0 * f : f;
}

struct CatPlan {
  bool ok = false;
  int64_t m, n, c, a, b;
  int64_t col_W, col_us;
  bool col_staged;
  int64_t row_us;
  bool row_staged;
  int npass;
  double cost;
};

// Catanzaro passes act on an m x n grid with m <= n (C2R on R x C when
// C >= R, R2C on the C x R view otherwise): columns of m units, rows of n.
CatPlan plan_catanzaro(int64_t R, int64_t C, int64_t U, const DeviceProfile& prof,
                       int64_t scratch_budget) {
  CatPlan p;
  p.m = std::min(R, C);
  p.n = std::max(R, C);
  p.c = gcd64(p.m, p.n);
  p.a = p.m / p.c;
  p.b = p.n / p.c;
  p.npass = p.c > 1 ? 3 : 2;
  // A pass whose footprint (`units` whole units) fits neither LDS at an
  // aligned slice: staged through scratch if one footprint fits the budget,
  // else narrow slices in LDS, else narrow slices through scratch.
  double narrow_cost = 0;
  auto fallback = [&](int64_t units, int64_t& us, bool& staged) {
    if (units * U <= scratch_budget) {
      us = U;
      staged = true;
    } else if ((us = any_slice(units, U, prof.lds_max))) {
      staged = false;
      narrow_cost += 0.5;
    } else if ((us = any_slice(units, U, scratch_budget))) {
      staged = true;
      narrow_cost += 0.5;
    } else {
      us = U;
      staged = true;  // footprint exceeds the budget: rejected below
    }
  };
  // column passes
  p.col_staged = false;
  if (U >= 128) {
    p.col_W = 1;
    p.col_us = pick_slice(p.m, U, prof.lds_max, prof.min_segment);
    if (!p.col_us) fallback(p.m, p.col_us, p.col_staged);
  } else {
    // panels of 512-byte row segments when they fit 1.25 x one workgroup's
    // LDS target (rows of a column pass lie a whole grid row apart, so longer
    // segments make fuller HBM bursts: 141 rows of 30-byte units took 2.22 ms
    // in 150-byte segments and 1.57 ms in 540-byte ones of 76 KB panels),
    // else of 128 bytes
    int64_t W = std::min<int64_t>(p.n, (512 + U - 1) / U);
    if (p.m * W * U > std::min(prof.lds_target + prof.lds_target / 4, prof.lds_max))
      W = std::min<int64_t>(p.n, (128 + U - 1) / U);
    if (p.m * W * U > prof.lds_max) W = std::min<int64_t>(p.n, (prof.min_segment + U - 1) / U);
    if (p.m * W * U > prof.lds_max) {
      W = std::min<int64_t>(p.n, (128 + U - 1) / U);
      p.col_staged = true;
    }
    p.col_W = W;
    p.col_us = U;
  }
  // row pass
  p.row_staged = false;
  p.row_us = pick_slice(p.n, U, prof.lds_max, prof.min_segment);
  if (!p.row_us) fallback(p.n, p.row_us, p.row_staged);
  const double col_cost = cat_pass_cost(p.col_us, p.col_staged) * (p.col_staged ? prof.staged_mult : 1.0),
               row_cost = cat_pass_cost(p.row_us, p.row_staged) * (p.row_staged ? prof.staged_mult : 1.0);
  p.cost = row_cost + col_cost * (p.npass - 1) + narrow_cost;
  // a staged pass needs at least one footprint of scratch
  const int64_t col_fp = p.m * p.col_W * p.col_us, row_fp = p.n * p.row_us;
  p.ok = true;
  if (p.col_staged && col_fp > scratch_budget) p.ok = false;
  if (p.row_staged && row_fp > scratch_budget) p.ok = false;
  // small penalty per staged pass so that equal-traffic LDS plans win ties
  p.cost += 0.01 * ((p.col_staged ? p.npass - 1 : 0) + (p.row_staged ? 1 : 0));
  return p;
}

const std::vector<int64_t>& divisors(int64_t x) {
  thread_local std::map<int64_t, std::vector<int64_t>> memo;
  auto it = memo.find(x);
  if (it != memo.end()) return it->second;
  if (memo.size() > 4096) memo.clear();
  std::vector<int64_t> d;
  for (int64_t q = 1; q * q <= x; ++q)
    if (x % q == 0) {
      d.push_back(q);
      if (q != x / q) d.push_back(x / q);
    }
  std::sort(d.begin(), d.end());
  return memo[x] = std::move(d);
}

// Segment length of a cycle pass over B regions of N cells of U bytes:
// short enough for about 32 Ki waves (one per segment and 1 KiB of cell),
// long enough that the saved boundary cells and the table fit the budget.
std::shared_ptr<CycleTable> choose_cycle_table(const std::vector<CycleMap>& maps, int64_t N,
                                               int64_t B, int64_t U, int64_t budget,
                                               int64_t lmax0) {
  const int64_t chunk = 64 * pow2_factor(U, 16);
  const int64_t chunks = (U + chunk - 1) / chunk;
  int64_t lmax = lmax0 > 0 ? lmax0 : 256;
  std::shared_ptr<CycleTable> t = build_cycle_table_maps(maps, N, lmax);
  while (!lmax0 && lmax > 16 && B * (int64_t)t->segs.size() * chunks < 32768) {
    lmax /= 2;
    t = build_cycle_table_maps(maps, N, lmax);
  }
  while (!t->slots.empty() && B * (int64_t)t->slots.size() * U + t->bytes() > budget) {
    lmax = lmax >= N ? N : lmax * 2;
    t = build_cycle_table_maps(maps, N, lmax);
  }
  return t;
}

enum How {
  HOW_NONE = 0,
  HOW_TILE = 1,
  HOW_CAT = 2,
  HOW_SPLIT_R = 3,
  HOW_SPLIT_C = 4,
  HOW_TILE_STAGED = 5,
  HOW_CYCLE = 6
};

// Relative time of one pass over the tensor, by pass class (contiguous LDS
// tile = 1, 5.7 TB/s of reads + writes on MI355X), from measurements on
// 4.5-172 GB tensors: sliced LDS tiles move short scattered segments (2.3x),
// staged tiles cross HBM twice (2.05-2.5x, 2.4x when sliced), cycle following
// reads and writes each cell once at scattered positions (1.6x, 1.75x for
// cells under 1 KiB); Catanzaro passes evaluate heavier index maps.
constexpr double kCostTile = 1.0;
constexpr double kCostTileSliced = 2.3;
constexpr double kCostTileStaged = 2.2;
constexpr double kCostTileStagedSliced = 2.4;
constexpr double kCostSplit = 0.001;

// Staged tiles of cells under 16 bytes whose output columns do not pack
// into 16-byte stores transpose word by word (about 0.7 TB/s measured).
inline double staged_tile_cost(int64_t R, int64_t U, int64_t us) {
  const double c = us == U ? kCostTileStaged : kCostTileStagedSliced;
  const int64_t w = pow2_factor(us, 16);
  const bool packs = us == w && R % (16 / w) == 0;
  return us < 16 && !packs ? 3.0 * c : c;
}

// Cycle-following moves whole cells (of at least cycle_min_cell bytes), one
// read and one write each; smaller cells waste HBM bursts.
constexpr uint64_t kCycleMaxCells = 1ull << 26;
// Composite cycle tables of two-pass regroup plans may span more cells: the
// plan replaces a whole pass of a large tensor (host walk: about 40 ns per
// cell, once per shape).
constexpr uint64_t kCycleMaxCellsRegroup = 1ull << 28;
constexpr double kCostCycle = 1.15;  // 4.6-5.8 TB/s for cells of 128 B to 384 KB, 16-byte words
constexpr double kCostBlock = 1.1;   // fused block permutation
// Cells whose size is not a multiple of 16 bytes: the single-launch kernel
// moves them in 16-byte chunks with a funnel shift (kCostCycleShift); the
// batched kernel moves narrower words (0.8-0.9 TB/s measured for 2- and
// 4-byte words).
constexpr double kCostCycleShift = 1.35;
thread_local bool g_cycle_word_cost = true;  // off: every cell width costs the same
thread_local bool g_cycle_cell_pref = false;  // on: larger cells cost less
thread_local bool g_cycle_shift = true;       // narrow cells move in shifted 16-byte chunks
inline double cycle_cost(int64_t U) {
  const int64_t w = pow2_factor(U, 16);
  if (!g_cycle_word_cost) return kCostCycle;
  if (w < 16) return g_cycle_shift ? kCostCycleShift : w == 8 ? 2.0 : w == 4 ? 4.0 : 6.0;
  if (!g_cycle_cell_pref) return kCostCycle;
  return U >= 512 ? kCostCycle - 0.1 : U >= 256 ? kCostCycle - 0.05 : kCostCycle;
}

// Widest word (at most 16 bytes) a segmented tile moves: it divides the
// segments, both row lengths and the unit, or the unit divides it.
int64_t segtile_word(const Op& o) {
  int64_t w = pow2_factor(gcd64(gcd64(o.segb, o.C * o.U), o.R * o.U), 16);
  while (w > 1 && (o.U % w) && (w % o.U)) w /= 2;
  return w;
}
// Segmented tile: a tile pass whose loads and stores are runs of segb bytes.
inline double segtile_cost(const Op& o) {
  const int64_t w = segtile_word(o);
  const double run = o.segb >= 256 ? 1.0 : o.segb >= 128 ? 1.1 : o.segb >= 64 ? 1.5 : 2.5;
  return run * (w >= 16 ? 1.0 : w == 8 ? 1.5 : 3.0);
}

struct Choice {
  int how = HOW_NONE;
  int64_t f = 0;   // split factor (the fast part)
  double cost = 1e30;
};

struct Lowerer {
  const DeviceProfile& prof;
  int64_t scratch_budget;
  std::map<std::tuple<int64_t, int64_t, int64_t, int>, Choice> memo;

  Lowerer(const DeviceProfile& p, int64_t budget) : prof(p), scratch_budget(budget) {}

  Choice best(int64_t R, int64_t C, int64_t U, int depth) {
    Choice ch;
    if (R <= 1 || C <= 1) {
      ch.how = HOW_NONE;
      ch.cost = 0;
      return ch;
    }
    auto key = std::make_tuple(R, C, U, 0);  // first result wins at any depth
    auto it = memo.find(key);
    if (it != memo.end()) return it->second;
    const int64_t mask = prof.lower_mask;
    if (int64_t us = (mask & LOWER_TILE) ? pick_tile_slice(R, C, U, prof) : 0) {
      ch.how = HOW_TILE;
      ch.cost = us == U ? kCostTile : kCostTileSliced;
      if (us == U) {
        memo[key] = ch;
        return ch;
      }
    }
    CatPlan cp = plan_catanzaro(R, C, U, prof, scratch_budget);
    if ((mask & LOWER_CATANZARO) && cp.ok && cp.cost < ch.cost) {
      ch.how = HOW_CAT;
      ch.cost = cp.cost;
    }
    // whole cells along the cycles of the transpose, no LDS
    if ((mask & LOWER_CYCLE) && prof.cycle_min_cell > 0 && U >= prof.cycle_min_cell &&
        (uint64_t)R * (uint64_t)C <= kCycleMaxCells &&
        cycle_cost(U) < ch.cost) {
      ch.how = HOW_CYCLE;
      ch.cost = cycle_cost(U);
    }
    // whole matrices (or slices) through scratch: copy out, tiled transpose back
    if ((mask & LOWER_STAGED_TILE) && ch.cost > kCostTileStaged * prof.staged_mult) {
      const int64_t us = pick_staged_tile_slice(R, C, U, scratch_budget, prof);
      const double cost = staged_tile_cost(R, U, us) * prof.staged_mult;
      if (us && cost < ch.cost) {
        ch.how = HOW_TILE_STAGED;
        ch.cost = cost;
      }
    }
    if ((mask & LOWER_SPLIT) && depth < 3 && ch.cost > 2 * kCostTile + 1e-9) {
      // R = R1 * R0: (B*R1, R0, C, U) then (B, R1, C, R0*U)
      for (int64_t r0 : divisors(R)) {
        if (r0 <= 1 || r0 >= R) continue;
        double c1 = best(r0, C, U, depth + 1).cost;
        if (c1 >= ch.cost) continue;
        double c2 = best(R / r0, C, r0 * U, depth + 1).cost;
        if (c1 + c2 + kCostSplit < ch.cost) {
          ch.how = HOW_SPLIT_R;
          ch.f = r0;
          ch.cost = c1 + c2 + kCostSplit;
        }
      }
      // C = C1 * C0: (B, R, C1, C0*U) then (B*C1, R, C0, U)
      for (int64_t c0 : divisors(C)) {
        if (c0 <= 1 || c0 >= C) continue;
        double c2 = best(R, c0, U, depth + 1).cost;
        if (c2 >= ch.cost) continue;
        double c1 = best(R, C / c0, c0 * U, depth + 1).cost;
        if (c1 + c2 + kCostSplit < ch.cost) {
          ch.how = HOW_SPLIT_C;
          ch.f = c0;
          ch.cost = c1 + c2 + kCostSplit;
        }
      }
    }
    memo[key] = ch;
    return ch;
  }


  void emit(int64_t B, int64_t R, int64_t C, int64_t U, uint64_t base, int stage, int depth,
            std::vector<Op>& ops) {
    if (R <= 1 || C <= 1 || B <= 0) return;
    Choice ch = best(R, C, U, depth);
    switch (ch.how) {
      case HOW_TILE: {
        Op op;
        op.kind = OP_TILE;
        op.stage = stage;
        op.base = base;
        op.B = B;
        op.R = R;
        op.C = C;
        op.U = U;
        op.us = pick_tile_slice(R, C, U, prof);
        ops.push_back(op);
        return;
      }
      case HOW_TILE_STAGED: {
        Op op;
        op.kind = OP_TILE;
        op.staged = 1;
        op.stage = stage;
        op.base = base;
        op.B = B;
        op.R = R;
        op.C = C;
        op.U = U;
        op.us = pick_staged_tile_slice(R, C, U, scratch_budget, prof);
        ops.push_back(op);
        return;
      }
      case HOW_CYCLE: {
        Op op;
        op.kind = OP_CYCLE;
        op.stage = stage;
        op.base = base;
        op.B = B;
        op.R = R;
        op.C = C;
        op.U = U;
        op.us = U;
        op.cmaps = {CycleMap{R * C, R, C, 1}};  // table built once the plan is chosen
        ops.push_back(op);
        return;
      }
      case HOW_CAT: {
        CatPlan cp = plan_catanzaro(R, C, U, prof, scratch_budget);
        const bool c2r = C >= R;
        Op g;
        g.stage = stage;
        g.base = base;
        g.B = B;
        g.R = cp.m;  // pass grid: m rows of n units
        g.C = cp.n;
        g.U = U;
        g.m = cp.m;
        g.n = cp.n;
        g.c = cp.c;
        g.a = cp.a;
        g.b = cp.b;
        g.k = modinv64(cp.a % cp.b, cp.b);
        g.binv = modinv64(cp.b % cp.a, cp.a);
        Op col = g, row = g;
        col.kind = OP_COL;
        col.W = cp.col_W;
        col.us = cp.col_us;
        col.staged = cp.col_staged;
        row.kind = OP_ROW;
        row.us = cp.row_us;
        row.staged = cp.row_staged;
        if (c2r) {
          if (cp.c > 1) {
            col.map = MAP_ROT;
            ops.push_back(col);
          }
          row.map = MAP_ROWG;
          ops.push_back(row);
          col.map = MAP_COLSH;
          ops.push_back(col);
        } else {
          col.map = MAP_COLSHINV;
          ops.push_back(col);
          row.map = MAP_ROWD;
          ops.push_back(row);
          if (cp.c > 1) {
            col.map = MAP_ROTINV;
            ops.push_back(col);
          }
        }
        return;
      }
      case HOW_SPLIT_R: {
        const int64_t r0 = ch.f, r1 = R / r0;
        emit(B * r1, r0, C, U, base, stage, depth + 1, ops);
        emit(B, r1, C, r0 * U, base, stage, depth + 1, ops);
        return;
      }
      case HOW_SPLIT_C: {
        const int64_t c0 = ch.f, c1 = C / c0;
        emit(B, R, c1, c0 * U, base, stage, depth + 1, ops);
        emit(B * c1, R, c0, U, base, stage, depth + 1, ops);
        return;
      }
      default: {
        Op bad;  // unreachable with a positive scratch budget; flagged by make_plan
        bad.kind = 0;
        bad.stage = stage;
        ops.push_back(bad);
        return;
      }
    }
  }
};

}  // namespace

// ---- tuning knobs --------------------------------------------------------------
namespace {
std::mutex g_tune_mu;
Tuning g_tune;
}  // namespace

Tuning tuning_get() {
  std::lock_guard<std::mutex> lk(g_tune_mu);
  return g_tune;
}

static int64_t* tuning_field(Tuning& t, const char* key) {
  if (!key) return nullptr;
  if (!strcmp(key, "lds_target")) return &t.lds_target;
  if (!strcmp(key, "tile_fp_target")) return &t.tile_fp_target;
  if (!strcmp(key, "slice_min")) return &t.slice_min;
  if (!strcmp(key, "stage_chunk")) return &t.stage_chunk;
  if (!strcmp(key, "cycle_min_cell")) return &t.cycle_min_cell;
  if (!strcmp(key, "fuse_max_block")) return &t.fuse_max_block;
  if (!strcmp(key, "block_digits")) return &t.block_digits;
  if (!strcmp(key, "block_pad")) return &t.block_pad;
  if (!strcmp(key, "cycle_word_cost")) return &t.cycle_word_cost;
  if (!strcmp(key, "threads")) return &t.threads;
  if (!strcmp(key, "threads_big")) return &t.threads_big;
  if (!strcmp(key, "xcd_remap")) return &t.xcd_remap;
  if (!strcmp(key, "cycle_variant")) return &t.cycle_variant;
  if (!strcmp(key, "reorder")) return &t.reorder;
  if (!strcmp(key, "regroup")) return &t.regroup;
  if (!strcmp(key, "cycle_lmax")) return &t.cycle_lmax;
  if (!strcmp(key, "cycle_cell_pref")) return &t.cycle_cell_pref;
  if (!strcmp(key, "segtile_threads")) return &t.segtile_threads;
  if (!strcmp(key, "lds_threads_auto")) return &t.lds_threads_auto;
  if (!strcmp(key, "segtile_nt")) return &t.segtile_nt;
  if (!strcmp(key, "lower_mask")) return &t.lower_mask;
  if (!strcmp(key, "hold_floor")) return &t.hold_floor;
  if (!strcmp(key, "cycle_u_depth")) return &t.cycle_u_depth;
  if (!strcmp(key, "min_groups")) return &t.min_groups;
  return nullptr;
}

int tuning_set(const char* key, int64_t value) {
  std::lock_guard<std::mutex> lk(g_tune_mu);
  Tuning t = g_tune;
  int64_t* f = tuning_field(t, key);
  if (!f || value < 0) return EITHOT_ERR_INVALID;
  const bool threads = f == &t.threads || f == &t.threads_big ||
                       (f == &t.segtile_threads && value != 0);
  if (threads && (value < 64 || value > 512 || value % 64)) return EITHOT_ERR_INVALID;
  if (f == &t.reorder && value > 3) return EITHOT_ERR_INVALID;
  if (f == &t.cycle_variant && value > 7) return EITHOT_ERR_INVALID;
  if (f == &t.cycle_lmax && value && (value < 2 || value > (1ll << 20))) return EITHOT_ERR_INVALID;
  if (f == &t.lds_target && value < 1024) return EITHOT_ERR_INVALID;
  if (f == &t.segtile_nt && value != 0 && value != 2 && value != 3) return EITHOT_ERR_INVALID;
  if (f == &t.lower_mask && (value == 0 || value > LOWER_ALL)) return EITHOT_ERR_INVALID;
  if (f == &t.cycle_u_depth && value != 4 && value != 6 && value != 8) return EITHOT_ERR_INVALID;
  if (f == &t.min_groups && (value < 1 || value > (1 << 20))) return EITHOT_ERR_INVALID;
  const bool flag = f == &t.regroup || f == &t.lds_threads_auto || f == &t.cycle_cell_pref ||
                    f == &t.cycle_word_cost || f == &t.block_digits || f == &t.block_pad ||
                    f == &t.xcd_remap;
  if (flag && value > 1) return EITHOT_ERR_INVALID;
  *f = value;
  t.epoch = g_tune.epoch + 1;
  g_tune = t;
  return EITHOT_OK;
}

int tuning_value(const char* key, int64_t* value) {
  std::lock_guard<std::mutex> lk(g_tune_mu);
  Tuning t = g_tune;
  int64_t* f = tuning_field(t, key);
  if (!f) return EITHOT_ERR_INVALID;
  if (value) *value = *f;
  return EITHOT_OK;
}

void tuning_apply(const Tuning& t, DeviceProfile& prof) {
  prof.lds_target = std::min<int64_t>(t.lds_target, prof.lds_max);
  prof.tile_fp_target = t.tile_fp_target;
  prof.slice_min = t.slice_min;
  prof.stage_chunk = t.stage_chunk;
  prof.cycle_min_cell = t.cycle_min_cell;
  prof.fuse_max_block = std::min<int64_t>(t.fuse_max_block, prof.lds_max);
  prof.block_digits = t.block_digits;
  prof.block_pad = t.block_pad;
  prof.cycle_word_cost = t.cycle_word_cost;
  prof.reorder = (int)t.reorder;
  prof.regroup = t.regroup;
  prof.cycle_lmax = t.cycle_lmax;
  prof.cycle_cell_pref = t.cycle_cell_pref;
  prof.lower_mask = t.lower_mask;
}

void lower_transpose2d(int64_t B, int64_t R, int64_t C, int64_t U, uint64_t base, int stage,
                       const DeviceProfile& prof, std::vector<Op>& ops) {
  Lowerer L(prof, INT64_MAX / 4);
  L.emit(B, R, C, U, base, stage, 0, ops);
}

static int64_t staged_scratch(const Op& op, int64_t budget, const DeviceProfile& prof) {
  if (op.kind == OP_CYCLE) return op.cyc ? op.B * (int64_t)op.cyc->slots.size() * op.U : 0;
  if (!op.staged) return 0;
  const int64_t fp = op.footprint_bytes();
  const int64_t nfp = op.num_footprints();
  if (prof.stage_chunk > 0) budget = std::min(budget, std::max(prof.stage_chunk, fp));
  int64_t g = std::max<int64_t>(1, budget / std::max<int64_t>(fp, 1));
  g = std::min(g, nfp);
  return g * fp;
}

// ---- fusion of nested contiguous tile passes -------------------------------------
// Consecutive whole-unit tile passes over the same bytes whose matrices
// (blocks of R*C*U bytes) nest form one permutation of every largest block;
// when that block fits `fuse_max_block`, one LDS pass applies it.
namespace {

bool fusable_tile(const Op& o) { return o.kind == OP_TILE && !o.staged && o.us == o.U; }

std::shared_ptr<BlockTable> compose_tiles(const std::vector<int64_t>& parts, int64_t F, int64_t g,
                                          bool digits, bool pad) {
  auto t = std::make_shared<BlockTable>();
  const int64_t n = F / g;
  t->cells = n;
  t->cell_bytes = g;
  std::vector<uint32_t> cur(n), nxt(n);
  for (int64_t i = 0; i < n; ++i) cur[i] = (uint32_t)i;
  for (size_t q = 0; q + 2 < parts.size(); q += 3) {
    const int64_t R = parts[q], C = parts[q + 1], k = parts[q + 2] / g, fb = R * C * k;
    for (int64_t b0 = 0; b0 < n; b0 += fb)
      for (int64_t j = 0; j < C; ++j)
        for (int64_t i = 0; i < R; ++i)
          for (int64_t x = 0; x < k; ++x)
            nxt[b0 + (j * R + i) * k + x] = cur[b0 + (i * C + j) * k + x];
    cur.swap(nxt);
  }
  t->src = std::move(cur);
  // digit form: greedy radices from the table, then checked on every cell
  {
    const std::vector<uint32_t>& T = t->src;
    int64_t step = 1;
    int nd = 0;
    bool ok = T[0] == 0;
    while (ok && step < n) {
      if (nd == 8) {
        ok = false;
        break;
      }
      const int64_t st = T[step];
      int64_t r = 1;
      while (step * (r + 1) <= n && (int64_t)T[step * r] == r * st) ++r;
      while (r > 1 && n % (step * r)) --r;
      if (r < 2) {
        ok = false;
        break;
      }
      t->radix[nd] = (uint32_t)r;
      t->stride[nd] = (uint32_t)st;
      ++nd;
      step *= r;
    }
    for (int64_t o = 0; ok && o < n; ++o) {
      int64_t q = o, v = 0;
      for (int k = 0; k < nd; ++k) {
        v += (q % t->radix[k]) * (int64_t)t->stride[k];
        q /= t->radix[k];
      }
      ok = v == (int64_t)T[o];
    }
    t->ndig = ok && digits ? nd : 0;
  }
  // LDS padding: the gather reads V = 16/g cells per thread (one word
  // each) across 64 lanes; pick the pad interval with the fewest bank
  // conflicts (64 banks of 4 bytes, and 32-lane halves on 32 banks)
  if (pad && g <= 16 && 64 % g == 0) {
    const int64_t V = g < 16 ? 16 / g : 1;
    auto cost = [&](int pl) {
      double tot = 0;
      int64_t cnt = 0;
      const int64_t chunks = n / V;
      for (int64_t t0 = 0; t0 < chunks && t0 < 16 * 64; t0 += 64)
        for (int64_t v = 0; v < V; ++v)
          for (int banks : {32, 64}) {
            for (int64_t h = 0; h < 64; h += banks) {
              int hist[64] = {0};
              int worst = 0;
              for (int64_t l = h; l < h + banks && t0 + l < chunks; ++l) {
                const int64_t x = (int64_t)t->src[(t0 + l) * V + v] * g;
                const int64_t a = pl ? x + (x >> pl) * 4 : x;
                worst = std::max(worst, ++hist[(a / 4) % banks]);
              }
              tot += worst;
              ++cnt;
            }
          }
      return cnt ? tot / cnt : 0.0;
    };
    double best = cost(0) * 0.999;
    for (int pl : {6, 7, 8, 9}) {  // 64..512 bytes between pads
      const double c = cost(pl);
      if (c < best) {
        best = c;
        t->pad_log2 = pl;
      }
    }
  }
  return t;
}

// Consecutive passes that permute whole cells of at least cycle_min_cell
// bytes over the same bytes (cycle passes and whole-unit tiles) compose into
// one cycle pass over cells of their common size, when at least one of them
// is a cycle pass.
void fuse_cycles(std::vector<Op>& ops, const DeviceProfile& prof) {
  if (prof.cycle_min_cell <= 0) return;
  auto cellperm = [&](const Op& o) {
    return (o.kind == OP_CYCLE && !o.cmaps.empty()) ||
           (o.kind == OP_TILE && !o.staged && o.us == o.U);
  };
  auto nmaps = [](const Op& o) { return o.kind == OP_CYCLE ? o.cmaps.size() : (size_t)1; };
  std::vector<Op> out;
  size_t i = 0;
  while (i < ops.size()) {
    size_t j = i, maps_n = 0;
    int64_t g = 0;
    int cycles = 0;
    while (j < ops.size() && cellperm(ops[j]) && maps_n + nmaps(ops[j]) <= 4 &&
           ops[j].base == ops[i].base && ops[j].bytes() == ops[i].bytes()) {
      const int64_t g2 = gcd64(g, ops[j].U);
      if (g2 < prof.cycle_min_cell || ops[j].bytes() / (uint64_t)g2 > kCycleMaxCells) break;
      g = g2;
      cycles += ops[j].kind == OP_CYCLE;
      maps_n += nmaps(ops[j]);
      ++j;
    }
    // composing must not cost more than the parts (a smaller common cell
    // can force narrower words)
    while (j - i >= 2) {
      double parts = 0;
      int64_t gg = 0;
      for (size_t q = i; q < j; ++q) {
        parts += ops[q].kind == OP_CYCLE ? cycle_cost(ops[q].U) : kCostTile;
        gg = gcd64(gg, ops[q].U);
      }
      if (cycle_cost(gg) <= parts) {
        g = gg;
        break;
      }
      --j;
    }
    cycles = 0;
    for (size_t q = i; q < j; ++q) cycles += ops[q].kind == OP_CYCLE;
    if (j - i < 2 || !cycles) {
      out.push_back(ops[i]);
      ++i;
      continue;
    }
    // every map re-expressed in g-byte cells, in pass order
    std::vector<CycleMap> maps;
    for (size_t q = i; q < j; ++q) {
      const Op& o = ops[q];
      const int64_t f = o.U / g;
      if (o.kind == OP_TILE)
        maps.push_back(CycleMap{o.R * o.C * f, o.R, o.C, f});
      else
        for (const CycleMap& M : o.cmaps) maps.push_back(CycleMap{M.bc * f, M.R, M.C, M.m * f});
    }
    Op c;
    c.kind = OP_CYCLE;
    c.stage = ops[i].stage;
    c.base = ops[i].base;
    c.B = 1;
    c.R = (int64_t)(ops[i].bytes() / (uint64_t)g);
    c.C = 1;
    c.U = g;
    c.us = g;
    c.cmaps = maps;
    out.push_back(c);
    i = j;
  }
  ops.swap(out);
}

// Relative time of a plan under the pass costs above.
double plan_cost(const Plan& plan) {
  double c = 0;
  for (const Op& o : plan.ops) {
    switch (o.kind) {
      case OP_TILE:
        c += o.staged ? staged_tile_cost(o.R, o.U, o.us)
                      : (o.us == o.U ? kCostTile : kCostTileSliced);
        break;
      case OP_COL:
      case OP_ROW: c += cat_pass_cost(o.us, o.staged != 0); break;
      case OP_CYCLE: c += cycle_cost(o.U); break;
      case OP_BLOCK: c += kCostBlock; break;
      case OP_SEGTILE: c += segtile_cost(o); break;
    }
  }
  return c;
}

void fuse_tiles(std::vector<Op>& ops, const DeviceProfile& prof) {
  if (prof.fuse_max_block <= 0) return;
  std::vector<Op> out;
  size_t i = 0;
  while (i < ops.size()) {
    size_t j = i + 1;
    int64_t F = 0, g = 0;
    if (fusable_tile(ops[i])) {
      F = ops[i].R * ops[i].C * ops[i].U;
      g = ops[i].U;
      const uint64_t total = ops[i].bytes();
      while (j < ops.size() && fusable_tile(ops[j]) && ops[j].base == ops[i].base &&
             ops[j].bytes() == total) {
        const int64_t Fj = ops[j].R * ops[j].C * ops[j].U;
        const int64_t Fn = std::max(F, Fj);
        // (the padded block, under 1/16 larger, must still fit LDS)
        if ((F % Fj && Fj % F) || Fn > prof.fuse_max_block || Fn + Fn / 16 + 16 > prof.lds_max)
          break;
        F = Fn;
        g = gcd64(g, ops[j].U);
        ++j;
      }
    }
    // a lone tile of small cells whose rows are not whole 16-byte chunks, or
    // whose output rows do not pack into 16-byte stores, moves words one by
    // one in the tile kernel; the block kernel loads and stores 16 bytes
    bool lone = false;
    if (j - i == 1 && fusable_tile(ops[i]) && F <= prof.fuse_max_block) {
      const Op& o = ops[i];
      const int64_t w = pow2_factor(o.U, 16);
      const bool rows16 = (o.C * o.U) % 16 == 0;
      const bool packs = o.U == w && o.R % (16 / w) == 0;
      lone = o.U < 16 && !(rows16 && packs);
    }
    if (j - i < 2 && !lone) {
      out.push_back(ops[i]);
      ++i;
      continue;
    }
    Op b;
    b.kind = OP_BLOCK;
    b.stage = ops[i].stage;
    b.base = ops[i].base;
    b.B = (int64_t)(ops[i].bytes() / (uint64_t)F);
    b.R = F / g;
    b.C = 1;
    b.U = g;
    b.us = g;
    for (size_t q = i; q < j; ++q) b.bparts.insert(b.bparts.end(), {ops[q].R, ops[q].C, ops[q].U});
    out.push_back(b);
    i = j;
  }
  ops.swap(out);
}

Stage make_stage(int64_t P, int64_t X, int64_t Y, int64_t S) {
  Stage s;
  s.P = P;
  s.X = X;
  s.Y = Y;
  s.S = S;
  int n = 0;
  if (P > 1) s.dims[n++] = P;
  s.dims[n++] = X;
  s.dims[n++] = Y;
  if (S > 1) s.dims[n++] = S;
  s.rank = n;
  const int xi = P > 1 ? 2 : 1;
  n = 0;
  if (P > 1) s.perm[n++] = 1;
  s.perm[n++] = xi + 1;
  s.perm[n++] = xi;
  if (S > 1) s.perm[n++] = xi + 2;
  return s;
}

// A transpose whose fastest dim moves, as three steps (the decomposer's
// stage order is one plan among many; this family often needs fewer passes):
//   1. cells of the input's leading dims X move so that the output's leading
//      dims Y follow X, the other dims Z after them  (X, Y, Z);
//   2. one batched 2-D transpose of the X-by-Y matrices  (Y, X, Z);
//   3. cells of Y move into the output order of X and Z.
// Steps 1 and 3 are whole-cell permutations: decomposed like any
// permutation, lowered, and composed into single cycle passes by
// fuse_cycles; step 2 is a tile pass when an X-by-Y matrix fits LDS.
struct Regroup {
  std::vector<int64_t> q;  // canonical dims, fastest first
  std::vector<int> p;      // canonical 0-based perm: output k = input p[k]
  int64_t E = 0;           // bytes of the canonical element

  // Drops unit dims, merges dims adjacent in both orders, folds a leading
  // dim that stays in place into the element.  False when nothing is left
  // to transpose.
  bool canonical(const int64_t* dims, const int* perm1, int rank, size_t elem) {
    std::vector<int> keep(rank, -1);
    std::vector<int64_t> d;
    for (int i = 0; i < rank; ++i)
      if (dims[i] > 1) {
        keep[i] = (int)d.size();
        d.push_back(dims[i]);
      }
    std::vector<int> pp;
    for (int k = 0; k < rank; ++k)
      if (keep[perm1[k] - 1] >= 0) pp.push_back(keep[perm1[k] - 1]);
    // merge runs: input dim a followed by a+1 in the output too
    const int n0 = (int)d.size();
    std::vector<int> head(n0, 1);  // 1 when input dim a starts a run
    for (int k = 0; k + 1 < (int)pp.size(); ++k)
      if (pp[k + 1] == pp[k] + 1) head[pp[k + 1]] = 0;
    std::vector<int> run(n0);
    q.clear();
    for (int a = 0; a < n0; ++a) {
      if (head[a]) q.push_back(1);
      run[a] = (int)q.size() - 1;
      q.back() *= d[a];
    }
    p.clear();
    for (int k = 0; k < (int)pp.size(); ++k)
      if (head[pp[k]]) p.push_back(run[pp[k]]);
    E = (int64_t)elem;
    if (!p.empty() && p[0] == 0) {
      E *= q[0];
      q.erase(q.begin());
      p.erase(p.begin());
      for (int& v : p) --v;
    }
    return q.size() >= 2;
  }
};

// Lowered ops of one whole-cell permutation (cell dims fastest first,
// 1-based perm), the cheaper decomposer order.
double lower_cellperm(Lowerer& L, const std::vector<int64_t>& cd, const std::vector<int>& perm1,
                      int64_t cell, int stage0, const DeviceProfile& prof, std::vector<Op>& ops,
                      std::vector<Stage>& stages) {
  double best = 1e300;
  for (int order : {(int)REORDER_PRE, (int)REORDER_POST}) {
    std::vector<Stage> st;
    if (!decompose(cd.data(), perm1.data(), (int)cd.size(), order, st)) return 1e300;
    std::vector<Op> o;
    for (size_t s = 0; s < st.size(); ++s)
      L.emit(st[s].S, st[s].Y, st[s].X, st[s].P * cell, 0, stage0 + (int)s, 0, o);
    bool ok = true;
    for (const Op& op : o) ok = ok && op.kind != 0;
    if (!ok) continue;
    fuse_cycles(o, prof);
    Plan tmp;
    tmp.ops = o;
    const double c = plan_cost(tmp);
    if (c < best - 1e-9) {
      best = c;
      ops = std::move(o);
      stages = std::move(st);
    }
  }
  return best;
}

// The X / Y / Z split of one regroup plan and its two cell permutations.
struct RegroupSplit {
  std::vector<int> Z;  // canonical dims of Z, fastest first
  int64_t px = 1, py = 1, pz = 1;
  std::vector<int64_t> c1, c3;  // cell dims of steps 1 and 3
  std::vector<int> t1, t3;      // their 1-based perms

  // X = the first i canonical input dims, Y = the first j output dims.
  bool init(const Regroup& g, int i, int j, bool z_out_order) {
    const int n = (int)g.q.size();
    std::vector<int> inX(n, 0), inY(n, 0);
    for (int a = 0; a < i; ++a) {
      inX[a] = 1;
      px *= g.q[a];
    }
    for (int k = 0; k < j; ++k) {
      if (inX[g.p[k]]) return false;
      inY[g.p[k]] = 1;
      py *= g.q[g.p[k]];
    }
    if (z_out_order) {
      for (int k = j; k < n; ++k)
        if (!inX[g.p[k]]) Z.push_back(g.p[k]);
    } else {
      for (int a = i; a < n; ++a)
        if (!inY[a]) Z.push_back(a);
    }
    for (int a : Z) pz *= g.q[a];
    // step 1: cell dims i..n-1 (input order) -> (Y in output order, Z)
    for (int a = i; a < n; ++a) c1.push_back(g.q[a]);
    for (int k = 0; k < j; ++k) t1.push_back(g.p[k] - i + 1);
    for (int a : Z) t1.push_back(a - i + 1);
    // step 3: cell dims (X in input order, Z) -> output order of the rest
    std::vector<int> l3;
    for (int a = 0; a < i; ++a) l3.push_back(a);
    l3.insert(l3.end(), Z.begin(), Z.end());
    for (int a : l3) c3.push_back(g.q[a]);
    for (int k = j; k < n; ++k)
      t3.push_back((int)(std::find(l3.begin(), l3.end(), g.p[k]) - l3.begin()) + 1);
    return true;
  }
};

// Best plan of the regroup family for (X = the first i canonical input
// dims, Y = the first j output dims); false when none lowers.
bool regroup_plan(const Regroup& g, int i, int j, bool z_out_order, size_t elem, Lowerer& L,
                  const DeviceProfile& prof, Plan& plan) {
  RegroupSplit sp;
  if (!sp.init(g, i, j, z_out_order)) return false;
  const int64_t px = sp.px, py = sp.py, pz = sp.pz;
  Plan out;
  std::vector<Op> o1, o3;
  std::vector<Stage> s1, s3;
  if (lower_cellperm(L, sp.c1, sp.t1, px * g.E, 0, prof, o1, s1) >= 1e300) return false;
  std::vector<Op> o2;
  L.emit(pz, py, px, g.E, 0, (int)s1.size(), 0, o2);
  for (const Op& op : o2)
    if (op.kind == 0) return false;
  if (lower_cellperm(L, sp.c3, sp.t3, py * g.E, (int)s1.size() + 1, prof, o3, s3) >= 1e300)
    return false;
  const int64_t ue = g.E / (int64_t)elem;  // canonical element in tensor elements
  for (const Stage& s : s1) out.stages.push_back(make_stage(s.P * px * ue, s.X, s.Y, s.S));
  out.stages.push_back(make_stage(ue, px, py, pz));
  for (const Stage& s : s3) out.stages.push_back(make_stage(s.P * py * ue, s.X, s.Y, s.S));
  out.ops = std::move(o1);
  out.ops.insert(out.ops.end(), o2.begin(), o2.end());
  out.ops.insert(out.ops.end(), o3.begin(), o3.end());
  fuse_cycles(out.ops, prof);
  fuse_tiles(out.ops, prof);
  plan.stages = std::move(out.stages);
  plan.ops = std::move(out.ops);
  return true;
}

// The decomposer order with fewer stages for a cell permutation.
bool fewest_stages(const std::vector<int64_t>& cd, const std::vector<int>& perm1,
                   std::vector<Stage>& st) {
  std::vector<Stage> pre, post;
  if (!decompose(cd.data(), perm1.data(), (int)cd.size(), REORDER_PRE, pre) ||
      !decompose(cd.data(), perm1.data(), (int)cd.size(), REORDER_POST, post))
    return false;
  st = post.size() < pre.size() ? post : pre;
  return true;
}

// Two-pass regroup: both cell permutations of the regroup plan composed into
// one cycle pass over cells of gcd(X bytes, Y bytes), and the X-by-Y
// matrices transposed where their runs lie.  Tile last: after the cycle pass
// the runs of a matrix are its output rows, at their output addresses; tile
// first: the runs are its input rows at their input addresses, and the
// cycle pass follows.
bool regroup2_plan(const Regroup& g, int i, int j, bool z_out_order, bool tile_last,
                   size_t elem, const DeviceProfile& prof, Plan& plan) {
  if (prof.cycle_min_cell <= 0) return false;
  RegroupSplit sp;
  if (!sp.init(g, i, j, z_out_order)) return false;
  const int n = (int)g.q.size();
  const int64_t Xb = sp.px * g.E, Yb = sp.py * g.E, gc = gcd64(Xb, Yb);
  const int64_t fp_max = prof.tile_fp_target > 0 ? prof.tile_fp_target : prof.lds_max / 2;
  if (gc < prof.cycle_min_cell || sp.px * Yb > fp_max || (int)sp.Z.size() > 6) return false;
  if ((tile_last ? i : j) > 6 || sp.pz >= (1ll << 31)) return false;
  {
    // padded LDS rows and the segment addresses must fit one workgroup
    const int64_t rs = tile_row_stride(Xb, pow2_factor(g.E, 16));
    const int64_t nseg = tile_last ? sp.px : sp.py;
    if ((sp.py * rs + 15) / 16 * 16 + nseg * 8 > prof.lds_max) return false;
  }
  std::vector<Stage> s1, s3;
  if (!fewest_stages(sp.c1, sp.t1, s1) || !fewest_stages(sp.c3, sp.t3, s3)) return false;
  if (s1.size() + s3.size() > (size_t)kMaxCycleMaps) return false;
  std::vector<CycleMap> maps;
  int64_t N = 1;
  for (int side = 0; side < 2; ++side)
    for (const Stage& st : side ? s3 : s1) {
      const int64_t f = st.P * (side ? Yb : Xb) / gc;
      maps.push_back(CycleMap{st.Y * st.X * f, st.Y, st.X, f});
      N = N / gcd64(N, maps.back().bc) * maps.back().bc;
    }
  uint64_t vol = 1;
  for (int64_t d : g.q) vol *= (uint64_t)d;
  const uint64_t cells = vol * (uint64_t)g.E / (uint64_t)gc;
  if ((uint64_t)N > kCycleMaxCellsRegroup || cells % (uint64_t)N) return false;

  // byte strides of the canonical dims in the input and output layouts
  std::vector<int64_t> ist(n), ost(n);
  int64_t acc = g.E;
  for (int a = 0; a < n; ++a) {
    ist[a] = acc;
    acc *= g.q[a];
  }
  acc = g.E;
  for (int k = 0; k < n; ++k) {
    ost[g.p[k]] = acc;
    acc *= g.q[g.p[k]];
  }
  Op t;
  t.kind = OP_SEGTILE;
  t.stage = (int)s1.size();
  t.B = sp.pz;
  t.R = sp.py;
  t.C = sp.px;
  t.U = g.E;
  t.us = g.E;
  t.segb = tile_last ? Yb : Xb;
  const std::vector<int64_t>& zst = tile_last ? ost : ist;
  if (tile_last)
    for (int a = 0; a < i; ++a) {
      t.sin_radix[t.sin_n] = g.q[a];
      t.sin_stride[t.sin_n++] = ost[a];
    }
  else
    for (int k = 0; k < j; ++k) {
      t.sin_radix[t.sin_n] = g.q[g.p[k]];
      t.sin_stride[t.sin_n++] = ist[g.p[k]];
    }
  for (int a : sp.Z) {
    t.sout_radix[t.sout_n] = g.q[a];
    t.sout_stride[t.sout_n++] = zst[a];
  }
  Plan out;
  if (!maps.empty()) {
    Op c;
    c.kind = OP_CYCLE;
    c.stage = tile_last ? 0 : (int)s1.size() + 1;
    c.B = (int64_t)(cells / (uint64_t)N);
    c.R = N;
    c.C = 1;
    c.U = gc;
    c.us = gc;
    c.cmaps = maps;
    out.ops.push_back(c);
  }
  out.ops.insert(tile_last ? out.ops.end() : out.ops.begin(), t);
  const int64_t ue = g.E / (int64_t)elem;
  for (const Stage& s : s1) out.stages.push_back(make_stage(s.P * sp.px * ue, s.X, s.Y, s.S));
  out.stages.push_back(make_stage(ue, sp.px, sp.py, sp.pz));
  for (const Stage& s : s3) out.stages.push_back(make_stage(s.P * sp.py * ue, s.X, s.Y, s.S));
  plan.stages = std::move(out.stages);
  plan.ops = std::move(out.ops);
  return true;
}

// g with canonical dim a split into (f, q[a] / f), the fast factor first in
// both orders.
Regroup split_dim(const Regroup& g, int a, int64_t f) {
  Regroup h;
  h.E = g.E;
  for (int x = 0; x < (int)g.q.size(); ++x) {
    if (x == a) {
      h.q.push_back(f);
      h.q.push_back(g.q[a] / f);
    } else {
      h.q.push_back(g.q[x]);
    }
  }
  for (int v : g.p) {
    if (v == a) {
      h.p.push_back(a);
      h.p.push_back(a + 1);
    } else {
      h.p.push_back(v > a ? v + 1 : v);
    }
  }
  return h;
}

// Two-pass regroup plans whose X-by-Y matrix is too large for LDS: the
// last dim of X and the last dim of Y are split, and X (Y) takes only their
// fast factors.  The splits giving the largest common cell, then the
// largest matrix, are tried.  best_cost: the cost to beat; on success the
// cheaper plan is in `plan`.
bool regroup2_split(const Regroup& g, int i, int j, bool z_out_order, bool tile_last,
                    size_t elem, const DeviceProfile& prof, double best_cost, Plan& plan) {
  if (prof.cycle_min_cell <= 0) return false;
  RegroupSplit sp;
  if (!sp.init(g, i, j, z_out_order)) return false;
  const int64_t fp_max = prof.tile_fp_target > 0 ? prof.tile_fp_target : prof.lds_max / 2;
  if (sp.px * sp.py * g.E <= fp_max) return false;  // no split needed
  const int a = i - 1, b = g.p[j - 1];
  const int64_t xo = sp.px / g.q[a], yo = sp.py / g.q[b];  // X, Y without their last dims
  struct Cand {
    int64_t f, h, gc, tile;
  };
  std::vector<Cand> cands;
  for (int64_t f : divisors(g.q[a])) {
    const int64_t Xb = xo * f * g.E;
    if (Xb < prof.cycle_min_cell || Xb * yo > fp_max) continue;
    for (int64_t h : divisors(g.q[b])) {
      const int64_t Yb = yo * h * g.E;
      if (Xb * (Yb / g.E) > fp_max) break;  // divisors ascend
      const int64_t gc = gcd64(Xb, Yb);
      if (gc < prof.cycle_min_cell || (f == g.q[a] && h == g.q[b])) continue;
      cands.push_back(Cand{f, h, gc, Xb * (Yb / g.E)});
    }
  }
  std::sort(cands.begin(), cands.end(), [](const Cand& x, const Cand& y) {
    return x.gc != y.gc ? x.gc > y.gc : x.tile > y.tile;
  });
  bool found = false;
  for (size_t c = 0; c < cands.size() && c < 4; ++c) {
    Regroup g1 = cands[c].f < g.q[a] ? split_dim(g, a, cands[c].f) : g;
    const int b1 = g1.p[j - 1];
    Regroup g2 = cands[c].h < g1.q[b1] ? split_dim(g1, b1, cands[c].h) : g1;
    Plan q;
    if (!regroup2_plan(g2, i, j, z_out_order, tile_last, elem, prof, q)) continue;
    const double cost = plan_cost(q);
    if (cost < best_cost - 1e-9) {
      best_cost = cost;
      plan = std::move(q);
      found = true;
    }
  }
  return found;
}

}  // namespace

// `reserve`: bytes of the alpha budget kept back for the pass tables.
static int make_plan_impl(const int64_t* dims_in, const int* perm1, int rank, size_t elem,
                          int n_sub, double alpha, const DeviceProfile& prof, int reorder,
                          int64_t reserve, Plan& plan, Lowerer* shared = nullptr) {
  plan = Plan();
  if (rank < 1 || rank > 64 || !dims_in || !perm1) return EITHOT_ERR_INVALID;
  if (elem == 0 || elem > (1u << 20)) return EITHOT_ERR_ELEM;
  if (!(alpha > 0) || !std::isfinite(alpha)) return EITHOT_ERR_INVALID;
  std::vector<int64_t> dims(dims_in, dims_in + rank);
  uint64_t vol = 1;
  for (int64_t d : dims) {
    if (d < 1) return EITHOT_ERR_INVALID;
    if (vol > (UINT64_MAX / 16) / (uint64_t)d) return EITHOT_ERR_INVALID;
    vol *= (uint64_t)d;
  }
  plan.tensor_bytes = vol * elem;
  std::vector<Stage> stages;
  if (!decompose(dims.data(), perm1, rank, reorder, stages)) return EITHOT_ERR_INVALID;
  plan.stages = stages;
  const int64_t budget = std::max<int64_t>(
      (int64_t)(alpha * (double)plan.tensor_bytes) - reserve, (int64_t)1);
  Lowerer own(prof, budget);
  // a caller's lowerer (same profile and budget) keeps its memo across orders
  Lowerer& L = shared && shared->scratch_budget == budget ? *shared : own;

  // Sub-tensor partitioning along the largest dim (Nd_transpose.cpp:265-300):
  // only when the caller asks for it and the dim exceeds floor(32/alpha).
  int nb = 1, p = 0;
  for (int i = 1; i < rank; ++i)
    if (dims[i] > dims[p]) p = i;
  const int64_t large = (int64_t)(32.0 * (1.0 / alpha));
  if (n_sub > 1 && dims[p] > large && !stages.empty())
    nb = (int)std::min<int64_t>(n_sub, dims[p]);
  plan.n_sub = nb;
  if (nb > 1) {
    // The split dim D becomes nb - 1 parts of q indices and a last part of
    // r (r == q when nb divides D).  The reference pads D up to a multiple
    // of nb instead, which grows the tensor (Nd_padding.cu:102-223); a
    // ragged last sub-tensor keeps the caller's buffer exactly the tensor.
    plan.partition_dim = p;
    const int64_t D = dims[p], E = (int64_t)elem;
    int64_t q = (D + nb - 1) / nb;
    if ((nb - 1) * q >= D) q = D / nb;
    const int64_t r = D - (nb - 1) * q;
    plan.part_len = q;
    plan.last_len = r;
    int64_t A = 1, S = 1;  // input: (A, D, S), fastest first
    for (int i = 0; i < p; ++i) A *= dims[i];
    for (int i = p + 1; i < rank; ++i) S *= dims[i];
    int pp = 0;  // output position of the split dim: (A2, D, S2)
    for (int i = 0; i < rank; ++i)
      if (perm1[i] - 1 == p) pp = i;
    int64_t A2 = 1, S2 = 1;
    for (int i = 0; i < pp; ++i) A2 *= dims[perm1[i] - 1];
    for (int i = pp + 1; i < rank; ++i) S2 *= dims[perm1[i] - 1];
    std::vector<int64_t> subq = dims, subr = dims;
    subq[p] = q;
    subr[p] = r;
    std::vector<Stage> sq, sr;
    decompose(subq.data(), perm1, rank, reorder, sq);
    decompose(subr.data(), perm1, rank, reorder, sr);
    plan.stages = sq;
    const int neq = r == q ? nb : nb - 1;  // sub-tensors of q indices
    if (r != q) {
      // Every slab s of the input is [X_s | Y_s] with X_s the first
      // (nb-1)*q indices of D and Y_s the last r: gather all X slabs before
      // all Y slabs.  In cells of g = gcd(X, Y) elements, that is the S x
      // (x+y) transpose followed by the x x S and y x S transposes of the
      // two regions it leaves.
      const int64_t X = A * q * (nb - 1), Y = A * r, g = gcd64(X, Y);
      const int64_t x = X / g, y = Y / g;
      L.emit(1, S, x + y, g * E, 0, -1, 0, plan.ops);
      L.emit(1, x, S, g * E, 0, -1, 0, plan.ops);
      L.emit(1, y, S, g * E, (uint64_t)(x * S * g * E), -1, 0, plan.ops);
    }
    // partition (Nd_decompose.cpp:16-25,48-56): (A*q, neq, S) -> 132, so
    // the neq sub-tensors of q indices are contiguous; the last one (when
    // ragged) already is, after them
    L.emit(1, S, neq, A * q * E, 0, -1, 0, plan.ops);
    // the sub-tensors of q indices share one batched plan
    for (size_t st = 0; st < sq.size(); ++st)
      L.emit(sq[st].S * neq, sq[st].Y, sq[st].X, sq[st].P * E, 0, (int)st, 0, plan.ops);
    const uint64_t last = (uint64_t)(A * q * neq * S * E);
    if (r != q)
      for (size_t st = 0; st < sr.size(); ++st)
        L.emit(sr[st].S, sr[st].Y, sr[st].X, sr[st].P * E, last, (int)st, 0, plan.ops);
    // join (Nd_decompose.cpp:32-41,65-82): (A2*q, S2, neq) -> 132
    L.emit(1, neq, S2, A2 * q * E, 0, -1, 0, plan.ops);
    if (r != q) {
      // the inverse of the gather above on the output side: slabs [X2_s]
      // then [Y2_s] back to [X2_s | Y2_s]
      const int64_t X2 = A2 * q * (nb - 1), Y2 = A2 * r, g2 = gcd64(X2, Y2);
      const int64_t x2 = X2 / g2, y2 = Y2 / g2;
      L.emit(1, S2, x2, g2 * E, 0, -1, 0, plan.ops);
      L.emit(1, S2, y2, g2 * E, (uint64_t)(x2 * S2 * g2 * E), -1, 0, plan.ops);
      L.emit(1, x2 + y2, S2, g2 * E, 0, -1, 0, plan.ops);
    }
  } else {
    for (size_t s = 0; s < stages.size(); ++s) {
      const Stage& st = stages[s];
      L.emit(st.S, st.Y, st.X, st.P * (int64_t)elem, 0, (int)s, 0, plan.ops);
    }
  }
  for (const Op& op : plan.ops)
    if (op.kind == 0) return EITHOT_ERR_ALPHA;
  fuse_cycles(plan.ops, prof);
  fuse_tiles(plan.ops, prof);
  return EITHOT_OK;
}

// Cycle tables of the chosen plan, then its scratch and table bytes.
static void finalize_plan(Plan& plan, double alpha, int64_t reserve, const DeviceProfile& prof) {
  const int64_t budget = std::max<int64_t>(
      (int64_t)(alpha * (double)plan.tensor_bytes) - reserve, (int64_t)1);
  plan.scratch_bytes = 0;
  plan.table_bytes = 0;
  for (Op& op : plan.ops) {
    if (op.kind == OP_BLOCK && !op.blk)
      op.blk = compose_tiles(op.bparts, op.R * op.U, op.U, prof.block_digits != 0, prof.block_pad != 0);
    if (op.kind == OP_CYCLE && !op.cyc) {
      op.cyc = choose_cycle_table(op.cmaps, op.R * op.C, op.B, op.U, budget, prof.cycle_lmax);
    }
    plan.scratch_bytes = std::max(plan.scratch_bytes, staged_scratch(op, budget, prof));
    if (op.cyc) plan.table_bytes += op.cyc->bytes();
    if (op.blk) plan.table_bytes += op.blk->bytes();
  }
}

static int make_plan_search(const int64_t* dims_in, const int* perm1, int rank, size_t elem,
                            int n_sub, double alpha, const DeviceProfile& prof, int reorder,
                            Plan& plan);

// Planning allocates host tables that can be large (a cycle table spans up
// to 2^28 cells); running out of host memory is reported, not thrown.
int make_plan(const int64_t* dims_in, const int* perm1, int rank, size_t elem, int n_sub,
              double alpha, const DeviceProfile& prof, int reorder, Plan& plan) {
  try {
    return make_plan_search(dims_in, perm1, rank, elem, n_sub, alpha, prof, reorder, plan);
  } catch (const std::bad_alloc&) {
    plan = Plan();
    return EITHOT_ERR_NOMEM;
  } catch (...) {
    plan = Plan();
    return EITHOT_ERR_INVALID;
  }
}

static int make_plan_search(const int64_t* dims_in, const int* perm1, int rank, size_t elem,
                            int n_sub, double alpha, const DeviceProfile& prof, int reorder,
                            Plan& plan) {
  g_cycle_word_cost = prof.cycle_word_cost != 0;
  g_cycle_cell_pref = prof.cycle_cell_pref != 0;
  g_cycle_shift = prof.cycle_shift != 0;
  // Search: decomposer order (pre/post, or the one asked for) x how hard
  // staged passes are avoided; cycle passes that later compose with their
  // neighbours make a split lowering cheaper than a staged one, which only
  // the whole-plan cost shows.
  std::vector<int> orders;
  if (reorder == 0 && rank > 0 && rank <= 64 && dims_in && perm1) {
    // the mix rule's order first, so that it wins ties; the other order only
    // when its stages differ
    std::vector<Stage> mix, pre, post;
    decompose(dims_in, perm1, rank, REORDER_MIX, mix);
    decompose(dims_in, perm1, rank, REORDER_PRE, pre);
    decompose(dims_in, perm1, rank, REORDER_POST, post);
    auto same = [](const std::vector<Stage>& a, const std::vector<Stage>& b) {
      if (a.size() != b.size()) return false;
      for (size_t i = 0; i < a.size(); ++i)
        if (a[i].P != b[i].P || a[i].X != b[i].X || a[i].Y != b[i].Y || a[i].S != b[i].S)
          return false;
      return true;
    };
    const bool mix_is_pre = same(mix, pre);
    orders = {mix_is_pre ? (int)REORDER_PRE : (int)REORDER_POST};
    if (!same(pre, post)) orders.push_back(mix_is_pre ? (int)REORDER_POST : (int)REORDER_PRE);
  } else {
    orders = {reorder == 3 ? (int)REORDER_MIX : reorder};
  }
  // one lowerer (and its memo) per staged weight, shared by the orders
  uint64_t vol = 1;
  bool vol_ok = rank > 0 && rank <= 64 && dims_in && elem > 0;
  for (int i = 0; vol_ok && i < rank; ++i) {
    if (dims_in[i] < 1 || vol > (UINT64_MAX / 16) / (uint64_t)dims_in[i]) vol_ok = false;
    else vol *= (uint64_t)dims_in[i];
  }
  const int64_t budget0 =
      vol_ok && alpha > 0 ? std::max<int64_t>((int64_t)(alpha * (double)(vol * elem)), 1) : 1;
  int rc = EITHOT_ERR_INVALID;
  double best_cost = 1e300;
  int best_order = orders[0];
  double best_mult = 1.0;
  // the regroup family, unless the caller asked for sub-tensor partitioning
  // and it applies
  bool partition = false;
  if (vol_ok && n_sub > 1 && alpha > 0) {
    int pd = 0;
    for (int i = 1; i < rank; ++i)
      if (dims_in[i] > dims_in[pd]) pd = i;
    partition = dims_in[pd] > (int64_t)(32.0 * (1.0 / alpha));
  }
  Regroup rg;
  const bool regroup = vol_ok && !partition && reorder == 0 && prof.regroup &&
                       rg.canonical(dims_in, perm1, rank, elem) && rg.q.size() >= 3;
  for (double mult : {1.0, 3.0}) {
    DeviceProfile p2 = prof;
    p2.staged_mult = mult;
    Lowerer shared(p2, budget0);
    for (int o : orders) {
      Plan q;
      const int r =
          make_plan_impl(dims_in, perm1, rank, elem, n_sub, alpha, p2, o, 0, q, &shared);
      if (r) {
        if (rc) rc = r;
        continue;
      }
      const double c = plan_cost(q);
      if (rc || c < best_cost - 1e-9) {
        best_cost = c;
        best_order = o;
        best_mult = mult;
        plan = std::move(q);
      }
      rc = EITHOT_OK;
    }
    if (!regroup || (rc && rc != EITHOT_ERR_ALPHA)) continue;
    const int n = (int)rg.q.size();
    for (int i = 1; i < n && i <= 6; ++i)
      for (int j = 1; i + j <= n && j <= 6; ++j)
        for (bool zo : {false, true}) {
          if (zo && i + j == n) continue;  // no Z: both orders agree
          for (int form = 0; form < 5; ++form) {
            Plan q;
            const bool ok =
                form == 0   ? regroup_plan(rg, i, j, zo, elem, shared, p2, q)
                : form <= 2 ? regroup2_plan(rg, i, j, zo, form == 1, elem, p2, q)
                            : regroup2_split(rg, i, j, zo, form == 3, elem, p2,
                                             rc ? 1e300 : best_cost, q);
            if (!ok) continue;
            q.tensor_bytes = vol * elem;
            const double c = plan_cost(q);
            if (rc || c < best_cost - 1e-9) {
              best_cost = c;
              best_mult = mult;
              plan = std::move(q);
              rc = EITHOT_OK;
            }
          }
        }
  }
  if (rc) return rc;
  DeviceProfile pb = prof;
  pb.staged_mult = best_mult;
  finalize_plan(plan, alpha, 0, pb);
  if (plan.table_bytes == 0) return EITHOT_OK;
  // The pass tables count as extra memory too.  When they push the plan
  // over its budget: plan again with their bytes kept back from the
  // scratch budget; then without cycle following; then without fusion;
  // keep whichever needs the least extra memory.
  const int64_t budget = (int64_t)(alpha * (double)plan.tensor_bytes);
  auto extra = [](const Plan& p) { return p.scratch_bytes + p.table_bytes; };
  if (extra(plan) <= budget) return EITHOT_OK;
  Plan best = plan;
  for (int attempt = 0; attempt < 3; ++attempt) {
    int64_t reserve = 0;
    if (attempt == 0) reserve = plan.table_bytes;
    if (attempt == 1) pb.cycle_min_cell = 0;
    if (attempt == 2) pb.fuse_max_block = 0;
    Plan q;
    if (make_plan_impl(dims_in, perm1, rank, elem, n_sub, alpha, pb, best_order, reserve, q))
      continue;
    finalize_plan(q, alpha, reserve, pb);
    if (extra(q) < extra(best)) best = q;
    if (extra(best) <= budget) break;
  }
  plan = best;
  return EITHOT_OK;
}

int64_t Plan::device_table_bytes(int dev) const {
  int64_t n = 0;
  std::vector<const void*> seen;  // a table shared by several passes counts once
  auto add = [&](const void* t, const DeviceCopies& c, int64_t bytes) {
    if (std::find(seen.begin(), seen.end(), t) != seen.end() || !c.has(dev)) return;
    seen.push_back(t);
    n += std::max<int64_t>(bytes, 16);  // the uploads allocate at least 16 bytes
  };
  for (const Op& op : ops) {
    if (op.cyc) add(op.cyc.get(), op.cyc->dev, op.cyc->bytes());
    if (op.blk) add(op.blk.get(), op.blk->dev, op.blk->bytes());
  }
  return n;
}

void Plan::drop_device_tables(int dev) const {
  for (const Op& op : ops) {
    if (op.cyc) op.cyc->dev.drop(dev);
    if (op.blk) op.blk->dev.drop(dev);
  }
}

std::string Plan::describe() const {
  std::ostringstream os;
  static const char* kn[] = {"?", "tile", "col", "row", "cycle", "block", "segtile"};
  static const char* mn[] = {"-", "rot", "rotinv", "colsh", "colshinv", "rowg", "rowd"};
  os << "stages:";
  for (const Stage& s : stages) {
    os << " ";
    for (int i = 0; i < s.rank; ++i) os << s.perm[i];
    os << "(";
    for (int i = 0; i < s.rank; ++i) os << (i ? "," : "") << s.dims[i];
    os << ")";
  }
  os << "\nn_sub=" << n_sub << " scratch=" << scratch_bytes << "\n";
  for (const Op& op : ops) {
    os << "  [" << op.stage << "] " << kn[op.kind] << "/" << mn[op.map] << " B=" << op.B
       << " R=" << op.R << " C=" << op.C << " U=" << op.U << " us=" << op.us;
    if (op.kind == OP_COL) os << " W=" << op.W;
    if (op.staged) os << " STAGED";
    os << "\n";
  }
  return os.str();
}

// ============================================================================
// Reference-parity estimator: the size model of the Nd_estimation twin
// (Nd_memory_estimation/src/{col_op.cu:15-176,row_op.cu:14-173,
// 2dtranspose.cpp,3dtranspose.cpp,4dtranspose.cpp:116-143,Nd_decompose.cpp,
// Nd_padding.cu:95-112,test_inplace.cpp:18-147}).  Only gmem-tier passes
// need scratch; every other tier returns 0.
// ============================================================================
namespace {

struct RefModel {
  const RefProfile& p;
  int64_t T;  // sizeof(T): 4 -> float, anything else -> double (transpose.cpp:10-18)

  static int msb(int64_t x) {
    int r = -1;
    while (x > 0) {
      ++r;
      x >>= 1;
    }
    return r;
  }
  int64_t nthr(bool rotate, int64_t d1) const {
    if (rotate) return 256;
    return msb(d1) <= 18 ? 32 : 256;
  }
  // col_op.cu:162-176 (_2d) and :111-128 (_213)
  int64_t col2d(bool rot, int64_t d1, int64_t d2) const {
    if (p.smem_per_block / (T * d2) >= 16) return 0;
    return T * nthr(rot, d1) * d2;
  }
  // col_op.cu:71-88 (_132)
  int64_t col132(bool rot, int64_t d1, int64_t d2) const {
    if (2 * T * d1 * d2 <= p.smem_per_block / 32) return 0;
    if (T * d1 * d2 <= p.smem_per_block) return 0;
    if (p.smem_per_block / (T * d2) >= 16) return 0;
    return T * nthr(rot, d1) * d2;
  }
  // row_op.cu:130-173 (_2d), :95-128 (_213)
  int64_t row2d(int64_t d1) const {
    if (2 * d1 * T <= p.smem_per_block / 32) return 0;
    if (T * d1 <= p.smem_per_block) return 0;
    return T * d1;
  }
  // row_op.cu:52-81 (_132)
  int64_t row132(int64_t d1, int64_t d3) const {
    if (T * d1 * d3 <= p.smem_per_block) return 0;
    return T * d1 * d3;
  }
  // 2dtranspose.cpp:21-94
  int64_t t2d(int64_t d1, int64_t d2) const {
    const int64_t c = gcd64(d1, d2);
    if (d1 >= d2)
      return std::max({c > 1 ? col2d(true, d1, d2) : 0, row2d(d1), col2d(false, d1, d2)});
    // r2c(d2, d1): the kernels see d2 columns and d1 rows
    return std::max({col2d(false, d2, d1), row2d(d2), c > 1 ? col2d(true, d2, d1) : 0});
  }
  // 3dtranspose.cpp:133-171 (213)
  int64_t t213(int64_t d1, int64_t d2, int64_t d3) const {
    if (d3 >= 8) return row2d(d1 * d2);
    const int64_t c = gcd64(d1, d2);
    if (d1 > d2)
      return std::max({c > 1 ? col2d(true, d1, d2) : 0, row2d(d1), col2d(false, d1, d2)});
    return std::max({col2d(false, d2, d1), row2d(d2), c > 1 ? col2d(true, d2, d1) : 0});
  }
  // 3dtranspose.cpp:175-228 (132)
  int64_t t132(int64_t d1, int64_t d2, int64_t d3) const {
    if (d1 >= 48) return col2d(false, d1, d2 * d3);
    if (d1 > 12) {
      if (d2 > d3) {  // c2r(d1, d3, d2): grid of d3 rows, d2 columns
        const int64_t c = gcd64(d3, d2);
        return std::max({c > 1 ? col132(true, d1, d3) : 0, row132(d1, d2), col132(false, d1, d3)});
      }
      const int64_t c = gcd64(d3, d2);  // r2c(d1, d2, d3)
      return std::max({col132(false, d1, d2), row132(d1, d3), c > 1 ? col132(true, d1, d2) : 0});
    }
    if (d2 >= d3) return std::max(t2d(d1 * d2, d3), row2d(d3 * d1));
    return std::max(row2d(d1 * d2), t2d(d2, d1 * d3));
  }
  // 4dtranspose.cpp:116-143 (the estimator's 1324 rule has no alpha term)
  int64_t t1324(int64_t d1, int64_t d2, int64_t d3, int64_t d4) const {
    if (d1 >= 96) return col2d(false, d1, d2 * d3 * d4);
    if (d4 >= 4) return row2d(d1 * d2 * d3);
    if (d2 >= d3) return std::max(t213(d1 * d2, d3, d4), t213(d3, d1, d2 * d4));
    return std::max({row2d(d1 * d2), t2d(d2, d1 * d3 * d4), col2d(false, d1 * d3, d4 * d2)});
  }
  int64_t stage(const Stage& s) const {
    const int64_t* x = s.dims;
    int code = 0;
    for (int i = 0; i < s.rank; ++i) code = code * 10 + s.perm[i];
    switch (code) {
      case 21: return t2d(x[0], x[1]);
      case 213: return t213(x[0], x[1], x[2]);
      case 132: return t132(x[0], x[1], x[2]);
      case 1324: return t1324(x[0], x[1], x[2], x[3]);
    }
    return 0;
  }
  // Nd_transpose.cpp (estimator twin):156-249, max over rounds
  int64_t series(const int64_t* dims, const int* perm1, int rank) const {
    std::vector<Stage> st;
    decompose(dims, perm1, rank, (int)tuning_get().reorder, st);
    int64_t m = 0;
    for (const Stage& s : st) m = std::max(m, stage(s));
    return m;
  }
};

}  // namespace

int estimate_reference(const int64_t* dims, const int* perm1, int rank, size_t elem,
                       double alpha, const RefProfile& prof, RefEstimate& out) {
  out = RefEstimate();
  if (rank < 1 || !dims || !perm1 || !(alpha > 0)) return EITHOT_ERR_INVALID;
  RefModel M{prof, elem == 4 ? 4 : 8};
  double vol = 1;
  for (int i = 0; i < rank; ++i) vol *= (double)dims[i];
  const double data = vol * (double)M.T;
  int64_t t0 = M.series(dims, perm1, rank);
  out.init_bytes = (double)t0;
  // test_inplace.cpp:41-43: nb = 2^ceil(log2(floor(T0/bytes/alpha))), 0 -> 1
  const int percent = (int)((double)t0 / data / alpha);
  int nb = percent > 0 ? (int)std::pow(2.0, std::ceil(std::log2((double)percent))) : 0;
  if (nb == 0) nb = 1;
  if (t0 > 0) {  // :46-49
    alpha = std::max(alpha, 1.0 / std::sqrt(vol / 32.0));
    out.new_ratio = alpha;
    out.printed_new_ratio = true;
  }
  const int64_t large = (int64_t)(32.0 / alpha);
  int mp = 0;  // Nd_memory_estimation/src/util.cpp:86-94 last argmax
  for (int i = 0; i < rank; ++i)
    if (dims[i] >= dims[mp]) mp = i;
  if (!(dims[mp] > large && nb > 1)) {  // :62-79
    out.nb = nb;
    out.alpha = std::ceil(alpha * 100.0) / 100.0;
    return EITHOT_OK;
  }
  out.decomposed = true;
  // the reference keeps these sizes in size_t, so halving truncates
  const int64_t D = dims[mp];
  const int64_t ivol = (int64_t)vol;
  int64_t transpose_tmp = t0 / nb;
  int64_t padding_tmp = 0;
  if (D % nb != 0) {
    const int64_t Dp = (D / nb + 1) * nb;
    padding_tmp = ivol / D * (Dp - D) * M.T;
    const int64_t pre = M.T * prof.padding_blocks_per_sm * prof.sms * 1024;
    padding_tmp = std::max(padding_tmp, pre);
  }
  // partition and join sizes (Nd_decompose.cpp:17-75 of the estimator)
  int64_t ds = 1, nds = 1;
  for (int i = 0; i <= mp; ++i) ds *= dims[i];
  for (int i = mp + 1; i < rank; ++i) nds *= dims[i];
  const int p132[3] = {1, 3, 2};
  int64_t pd[3] = {ds / nb, nb, nds};
  int64_t decompose_tmp = M.series(pd, p132, 3);
  int pp = 0;
  for (int i = 0; i < rank; ++i)
    if (perm1[i] - 1 == mp) pp = i;
  int64_t ds2 = 1, nds2 = 1;
  for (int i = 0; i <= pp; ++i) ds2 *= dims[perm1[i] - 1];
  for (int i = pp + 1; i < rank; ++i) nds2 *= dims[perm1[i] - 1];
  int64_t jd[3] = {ds2 / nb, nds2, nb};
  decompose_tmp = std::max(decompose_tmp, M.series(jd, p132, 3));
  int64_t final_tmp = padding_tmp + std::max(decompose_tmp, transpose_tmp);
  double ratio = (double)final_tmp / data;
  if (ratio > alpha) {  // :97-122 halve nb while that does not increase the total
    int64_t new_final = final_tmp;
    int new_nb = nb;
    while (final_tmp >= new_final && nb > 1) {
      nb = new_nb;
      final_tmp = new_final;
      new_nb /= 2;
      // the reference grows the dim even when it divides evenly (:107-108)
      const int64_t Dp = (D / new_nb + 1) * new_nb;
      padding_tmp = ivol / D * (Dp - D) * M.T;
      transpose_tmp *= 2;
      decompose_tmp /= 2;
      new_final = padding_tmp + std::max(decompose_tmp, transpose_tmp);
    }
    ratio = (double)final_tmp / data;
    ratio = std::ceil(ratio * 100.0) / 100.0;
    alpha = std::max(ratio, alpha);
  }
  out.nb = nb;
  out.final_bytes = (double)final_tmp;
  out.final_ratio = ratio;
  out.alpha = alpha;
  return EITHOT_OK;
}

int estimate_native(const int64_t* dims, const int* perm1, int rank, size_t elem, double alpha,
                    const DeviceProfile& prof, int* n_sub, double* alpha_out,
                    size_t* extra_bytes) {
  Plan plan;
  int rc = make_plan(dims, perm1, rank, elem, 1, alpha, prof, prof.reorder, plan);
  if (rc != EITHOT_OK) return rc;
  if (n_sub) *n_sub = plan.n_sub;
  if (extra_bytes) *extra_bytes = (size_t)(plan.scratch_bytes + plan.table_bytes);
  if (alpha_out)
    *alpha_out = plan.tensor_bytes ? (double)(plan.scratch_bytes + plan.table_bytes) / (double)plan.tensor_bytes : 0.0;
  return EITHOT_OK;
}

}  // namespace eithot
