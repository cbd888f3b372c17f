// TEST INFRASTRUCTURE ONLY (oracle). Not part of the shipped product path.
//
// Indexed binary min-heap with the exact tie and capacity behaviour of the
// reference's branch-and-bound queue (helper/minheap.cpp:16-277,
// helper/minheap.hpp:17-64). The branch-and-bound order of the reference
// depends on these details (equal keys: a newly pushed entry bubbles above
// older ones because bubbleUp swaps on `>=`; holes left by pop are tracked
// through minEmpty/maxFull; keys outside [0, maxSize) or already present are
// ignored by push), so the steps are kept alike.
#pragma once
#include <limits>
#include <utility>
#include <vector>

namespace oracle {

class BnbHeap {
 public:
  explicit BnbHeap(long n) : slots_(n, {-1.0, -1}), where_(n, -1), cap_(n) {}

  bool empty() const { return count_ == 0; }
  long size() const { return count_; }
  // helper/minheap.hpp:31-33: +inf when empty.
  double peak() const { return count_ == 0 ? std::numeric_limits<double>::infinity() : slots_[0].first; }
  // helper/minheap.cpp:77-82: -1 when the key is absent.
  double key_value(long key) const {
    if (key < 0 || key >= cap_ || where_[key] == -1) return -1.0;
    return slots_[where_[key]].first;
  }

  bool contains(long key) const { return key >= 0 && key < cap_ && where_[key] != -1; }

  // helper/minheap.cpp:137-159
  void push(double value, long key) {
    if (key < 0 || key >= cap_ || where_[key] != -1 || count_ == cap_) return;
    slots_[first_free_] = {value, key};
    where_[key] = first_free_;
    sift_up(first_free_);
    ++count_;
    if (first_free_ == last_used_ + 1) ++last_used_;
    if (count_ == cap_) {
      first_free_ = count_ + 1;
      return;
    }
    while (slots_[first_free_].second > -1) ++first_free_;
  }

  // helper/minheap.cpp:87-116
  std::pair<double, long> pop() {
    if (count_ == 0) return {-1.0, -1};
    const std::pair<double, long> top = slots_[0];
    slots_[0].first = std::numeric_limits<double>::infinity();
    swap_slots(0, last_used_);
    slots_[last_used_] = {-1.0, -1};
    where_[top.second] = -1;
    long landed = sift_down(0);
    if (last_used_ < first_free_) first_free_ = last_used_;
    while (last_used_ >= 0 && slots_[last_used_].second == -1) --last_used_;
    if (landed > last_used_) last_used_ = landed;
    --count_;
    return top;
  }

  // helper/minheap.cpp:37-52
  bool decrease(long key, double value) {
    if (key < 0 || key >= cap_) return false;
    const long at = where_[key];
    if (at != -1 && value < slots_[at].first) {
      slots_[at].first = value;
      sift_up(at);
      if (slots_[at].first != value) return true;
    }
    return false;
  }

 private:
  void swap_slots(long i, long j) {
    if (i == j || i < 0 || j < 0 || i >= cap_ || j >= cap_ || slots_[i].second == -1 || slots_[j].second == -1) return;
    std::swap(slots_[i], slots_[j]);
    where_[slots_[j].second] = j;
    where_[slots_[i].second] = i;
  }
  // helper/minheap.cpp:168-231 (children are examined up to the capacity, holes skipped)
  long sift_down(long i) {
    if (i < 0 || i >= cap_) return -1;
    long l = 2 * i + 1, r = 2 * i + 2;
    while ((l < cap_ || r < cap_) && (l != -1 || r != -1)) {
      const bool lok = l != -1 && l < cap_ && slots_[l].second != -1 && slots_[l].first <= slots_[i].first;
      const bool rok = r != -1 && r < cap_ && slots_[r].second != -1 && slots_[r].first <= slots_[i].first;
      long go = -1;
      if (lok && rok) go = (slots_[r].first <= slots_[l].first) ? r : l;
      else if (lok) go = l;
      else if (rok) go = r;
      if (go < 0) break;
      std::swap(slots_[go], slots_[i]);
      where_[slots_[i].second] = i;
      where_[slots_[go].second] = go;
      i = go;
      l = 2 * i + 1;
      r = 2 * i + 2;
    }
    return i;
  }
  // helper/minheap.cpp:235-262 (`>=`: equal keys move up)
  long sift_up(long i) {
    if (i < 0 || i >= cap_) return -1;
    while (i > 0) {
      const long p = (i - 1) / 2;
      if (!(slots_[p].first >= slots_[i].first)) break;
      std::swap(slots_[p], slots_[i]);
      where_[slots_[i].second] = i;
      where_[slots_[p].second] = p;
      i = p;
    }
    return i;
  }

  std::vector<std::pair<double, long>> slots_;
  std::vector<long> where_;
  long cap_;
  long count_ = 0;
  long first_free_ = 0;
  long last_used_ = -1;
};

}  // namespace oracle
