// Host emulation of the cycle-following pass (OP_CYCLE) over the planner's
// cycle tables (eithot_amd/csrc/plan.cpp: build_cycle_table), for the CPU
// tests.  Items (matrix, segment, chunk of the cell) run the device
// kernel's exact step order: D loads of the cells that follow, then D
// stores; items are scheduled sequentially, reversed, shuffled or
// interleaved one step at a time to model concurrent waves.
#include <algorithm>
#include <cstdint>
#include <cstring>
#include <random>
#include <vector>

#include "../eithot_amd/csrc/plan.hpp"

using eithot::CycleSeg;
using eithot::CycleTable;

namespace {

constexpr int D = 8;

struct Item {
  uint64_t b, s, off, len;  // off: byte offset inside the cell, len: bytes this item moves
  uint32_t p = 0, t = 0;
  std::vector<uint8_t> temp;
  bool done = false;
};

struct Ctx {
  uint8_t* buf;
  const CycleTable* tab;
  uint64_t R, C, U, RCU;
  std::vector<uint8_t> boundary;
  uint8_t* cell(uint64_t b, uint32_t p, uint64_t off) { return buf + b * RCU + (uint64_t)p * U + off; }
  uint32_t src(uint32_t p) const {
    uint64_t q = p;
    for (size_t k = tab->maps.size(); k-- > 0;) {
      const eithot::CycleMap& M = tab->maps[k];
      const uint64_t b = q / M.bc, r = q - b * M.bc, c = r / M.m, x = r - c * M.m;
      const uint64_t j = c / M.R, i = c - j * M.R;
      q = b * M.bc + (i * M.C + j) * M.m + x;
    }
    return (uint32_t)q;
  }

  void start(Item& it) {
    const CycleSeg& sg = tab->segs[it.s];
    it.p = sg.start;
    it.t = 0;
    it.temp.resize(it.len);
    if (sg.src_slot < 0)
      memcpy(it.temp.data(), cell(it.b, sg.start, it.off), it.len);
    else
      memcpy(it.temp.data(),
             boundary.data() + (it.b * tab->slots.size() + (uint64_t)sg.src_slot) * U + it.off, it.len);
  }
  // one batch of the kernel loop
  void step(Item& it) {
    const CycleSeg& sg = tab->segs[it.s];
    std::vector<uint8_t> v[D];
    uint32_t q[D];
    uint32_t pp = it.p;
    for (int k = 0; k < D; ++k) q[k] = pp = src(pp);
    for (int k = 0; k < D; ++k)
      if (it.t + k < sg.len) {
        v[k].resize(it.len);
        if (it.t + k + 1 == sg.len)
          v[k] = it.temp;
        else
          memcpy(v[k].data(), cell(it.b, q[k], it.off), it.len);
      }
    uint32_t dst = it.p;
    for (int k = 0; k < D; ++k)
      if (it.t + k < sg.len) {
        memcpy(cell(it.b, dst, it.off), v[k].data(), it.len);
        dst = q[k];
      }
    it.p = q[D - 1];
    it.t += D;
    if (it.t >= sg.len) it.done = true;
  }
};

}  // namespace

namespace {
int64_t run_table(std::shared_ptr<CycleTable> tab, int64_t B, int64_t R, int64_t C, int64_t U,
                  int64_t chunk, int mode, uint64_t seed, uint8_t* buf, int64_t* nslots_out);
}

extern "C" {

// Transposes B matrices of R x C cells of U bytes in buf; mode 0 sequential,
// 1 reversed, 2 shuffled, 3 interleaved.  Returns segments, or -1.
int64_t ch_run(int64_t B, int64_t R, int64_t C, int64_t U, int64_t lmax, int64_t chunk, int mode,
               uint64_t seed, uint8_t* buf, int64_t* nslots_out) {
  return run_table(eithot::build_cycle_table(R, C, lmax), B, R, C, U, chunk, mode, seed, buf,
                   nslots_out);
}

// The composition of nmaps cell maps (bc, R, C, m each) over N cells of U
// bytes, in one cycle pass.
int64_t ch_run_maps(int nmaps, const int64_t* maps, int64_t N, int64_t U, int64_t lmax,
                    int64_t chunk, int mode, uint64_t seed, uint8_t* buf) {
  std::vector<eithot::CycleMap> mv;
  for (int k = 0; k < nmaps; ++k)
    mv.push_back(eithot::CycleMap{maps[4 * k], maps[4 * k + 1], maps[4 * k + 2], maps[4 * k + 3]});
  return run_table(eithot::build_cycle_table_maps(mv, N, lmax), 1, N, 1, U, chunk, mode, seed, buf,
                   nullptr);
}

}  // extern "C"

namespace {
int64_t run_table(std::shared_ptr<CycleTable> tab, int64_t B, int64_t R, int64_t C, int64_t U,
                  int64_t chunk, int mode, uint64_t seed, uint8_t* buf, int64_t* nslots_out) {
  Ctx cx{buf, tab.get(), (uint64_t)R, (uint64_t)C, (uint64_t)U, (uint64_t)(R * C * U), {}};
  // boundary save pass
  const uint64_t ns = tab->slots.size();
  cx.boundary.resize(B * ns * U);
  for (int64_t b = 0; b < B; ++b)
    for (uint64_t s = 0; s < ns; ++s)
      memcpy(cx.boundary.data() + (b * ns + s) * U, cx.cell(b, tab->slots[s], 0), U);
  std::vector<Item> items;
  for (int64_t b = 0; b < B; ++b)
    for (size_t s = 0; s < tab->segs.size(); ++s)
      for (int64_t off = 0; off < U; off += chunk) {
        Item it;
        it.b = b;
        it.s = s;
        it.off = off;
        it.len = std::min<int64_t>(chunk, U - off);
        items.push_back(it);
      }
  std::mt19937_64 rng(seed);
  if (mode == 1) std::reverse(items.begin(), items.end());
  if (mode == 2 || mode == 3) std::shuffle(items.begin(), items.end(), rng);
  if (mode == 3) {
    for (Item& it : items) cx.start(it);
    bool any = true;
    while (any) {
      any = false;
      for (Item& it : items)
        if (!it.done) {
          cx.step(it);
          any = true;
        }
    }
  } else {
    for (Item& it : items) {
      cx.start(it);
      while (!it.done) cx.step(it);
    }
  }
  if (nslots_out) *nslots_out = (int64_t)ns;
  return (int64_t)tab->segs.size();
}
}  // namespace

extern "C" {

// Every position that is not a fixed point is written by exactly one
// segment; segments never exceed lmax.  Returns 0 when the table is sound.
int ch_check(int64_t R, int64_t C, int64_t lmax) {
  auto tab = eithot::build_cycle_table(R, C, lmax);
  const uint64_t N = (uint64_t)R * C;
  std::vector<int> hits(N, 0);
  auto src = [&](uint64_t p) { return (p % R) * C + p / R; };
  for (const CycleSeg& sg : tab->segs) {
    if ((int64_t)sg.len > lmax) return 1;
    uint64_t p = sg.start, prev = p;
    for (uint32_t t = 0; t < sg.len; ++t) {
      if (p >= N) return 2;
      ++hits[p];
      prev = p;
      p = src(p);
    }
    // a whole cycle records its last position, which maps back to its start
    if (sg.src_slot < 0 && (sg.last != prev || src(sg.last) != sg.start)) return 4;
  }
  for (uint64_t p = 0; p < N; ++p) {
    const bool fixed = src(p) == p;
    if (hits[p] != (fixed ? 0 : 1)) return 3;
  }
  return 0;
}

}  // extern "C"
