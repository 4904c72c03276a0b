// Least-recently-used cache of shared, immutable values (plans) keyed by
// string.  Eviction only drops the cache's own reference: a caller that
// still holds a value (a queued server request, a transpose between planning
// and launch) keeps it alive, and the value is destroyed when its last
// holder lets go.  Not thread-safe by itself; the owner locks around it.
// Header-only and free of HIP so that host tests can build it alone.
#pragma once

#include <cstddef>
#include <cstdint>
#include <functional>
#include <list>
#include <memory>
#include <string>
#include <unordered_map>
#include <utility>
#include <vector>

namespace eithot {

template <typename T>
class LruCache {
 public:
  using Ptr = std::shared_ptr<const T>;

  // capacity: most entries held; weight_cap: most total weight held (the
  // newest entry is kept even when it alone exceeds it)
  explicit LruCache(size_t capacity, uint64_t weight_cap = UINT64_MAX)
      : cap_(capacity ? capacity : 1), wcap_(weight_cap) {}

  // The cached value, marked most recently used; null when absent.
  Ptr get(const std::string& key) {
    auto it = index_.find(key);
    if (it == index_.end()) return nullptr;
    order_.splice(order_.begin(), order_, it->second);
    return it->second->value;
  }

  // Inserts (or replaces) key and evicts least recently used entries beyond
  // the capacity.  Returns the evicted values (the caller decides when they
  // may be destroyed, e.g. outside its lock).
  std::vector<Ptr> put(const std::string& key, Ptr value, uint64_t weight = 0) {
    std::vector<Ptr> evicted;
    auto it = index_.find(key);
    if (it != index_.end()) {
      evicted.push_back(std::move(it->second->value));
      weight_ -= it->second->weight;
      it->second->value = std::move(value);
      it->second->weight = weight;
      order_.splice(order_.begin(), order_, it->second);
    } else {
      order_.push_front(Entry{key, std::move(value), weight});
      index_[key] = order_.begin();
    }
    weight_ += weight;
    while (order_.size() > cap_ || (order_.size() > 1 && weight_ > wcap_)) {
      evicted.push_back(std::move(order_.back().value));
      weight_ -= order_.back().weight;
      index_.erase(order_.back().key);
      order_.pop_back();
    }
    return evicted;
  }

  // Removes every entry; returns the values that were held.
  std::vector<Ptr> clear() {
    std::vector<Ptr> out;
    for (auto& e : order_) out.push_back(std::move(e.value));
    order_.clear();
    index_.clear();
    weight_ = 0;
    return out;
  }

  // Visits entries from most to least recently used.
  void for_each(const std::function<void(const std::string&, const Ptr&)>& fn) const {
    for (const auto& e : order_) fn(e.key, e.value);
  }

  size_t size() const { return order_.size(); }
  size_t capacity() const { return cap_; }
  uint64_t weight() const { return weight_; }

 private:
  struct Entry {
    std::string key;
    Ptr value;
    uint64_t weight;
  };
  size_t cap_;
  uint64_t wcap_;
  uint64_t weight_ = 0;
  std::list<Entry> order_;  // front: most recently used
  std::unordered_map<std::string, typename std::list<Entry>::iterator> index_;
};

}  // namespace eithot
