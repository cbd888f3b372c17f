// Branch-and-bound queue of the device-resident solver driver.
//
// The same indexed binary min-heap as the host driver's (bnb_heap.h), with
// the reference's tie and capacity behaviour (helper/minheap.cpp:16-277:
// bubbleUp swaps on `>=`, children are examined up to the capacity with
// holes skipped, push ignores keys outside [0, capacity) or already present),
// stored in global memory and driven by one thread of the workgroup.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pgc {

struct DevHeap {
  double* key;   // [cap] slot value (-1 when free)
  int32_t* val;  // [cap] slot key (branch id, -1 when free)
  int32_t* where;  // [cap] slot of a branch id (-1 when absent)
  int32_t cap;
  int32_t count, first_free, last_used;

  __device__ void init(double* k, int32_t* v, int32_t* w, int32_t n) {
    key = k;
    val = v;
    where = w;
    cap = n;
    count = 0;
    first_free = 0;
    last_used = -1;
  }
  __device__ bool empty() const { return count == 0; }
  __device__ double peak() const { return count == 0 ? __builtin_huge_val() : key[0]; }

  __device__ void swap_raw(int i, int j) {
    const double tk = key[i];
    const int32_t tv = val[i];
    key[i] = key[j];
    val[i] = val[j];
    key[j] = tk;
    val[j] = tv;
  }
  __device__ void swap_slots(int i, int j) {
    if (i == j || i < 0 || j < 0 || i >= cap || j >= cap || val[i] == -1 || val[j] == -1) return;
    swap_raw(i, j);
    where[val[j]] = j;
    where[val[i]] = i;
  }
  __device__ int sift_up(int i) {
    if (i < 0 || i >= cap) return -1;
    while (i > 0) {
      const int p = (i - 1) / 2;
      if (!(key[p] >= key[i])) break;
      swap_raw(p, i);
      where[val[i]] = i;
      where[val[p]] = p;
      i = p;
    }
    return i;
  }
  __device__ int sift_down(int i) {
    if (i < 0 || i >= cap) return -1;
    int l = 2 * i + 1, r = 2 * i + 2;
    while (l < cap || r < cap) {
      const bool lok = l < cap && val[l] != -1 && key[l] <= key[i];
      const bool rok = r < cap && val[r] != -1 && key[r] <= key[i];
      int go = -1;
      if (lok && rok) go = (key[r] <= key[l]) ? r : l;
      else if (lok) go = l;
      else if (rok) go = r;
      if (go < 0) break;
      swap_raw(go, i);
      where[val[i]] = i;
      where[val[go]] = go;
      i = go;
      l = 2 * i + 1;
      r = 2 * i + 2;
    }
    return i;
  }
  __device__ void push(double value, int k) {
    if (k < 0 || k >= cap || where[k] != -1 || count == cap || first_free >= cap) return;
    key[first_free] = value;
    val[first_free] = k;
    where[k] = first_free;
    sift_up(first_free);
    ++count;
    if (first_free == last_used + 1) ++last_used;
    if (count == cap) {
      first_free = count + 1;
      return;
    }
    while (first_free < cap && val[first_free] > -1) ++first_free;
  }
  // returns the popped branch id (its value in *v), -1 when empty
  __device__ int pop(double* v) {
    if (count == 0) {
      *v = -1.0;
      return -1;
    }
    const double tk = key[0];
    const int tv = val[0];
    key[0] = __builtin_huge_val();
    swap_slots(0, last_used);
    key[last_used] = -1.0;
    val[last_used] = -1;
    where[tv] = -1;
    const int landed = sift_down(0);
    if (last_used < first_free) first_free = last_used;
    while (last_used >= 0 && val[last_used] == -1) --last_used;
    if (landed > last_used) last_used = landed;
    --count;
    *v = tk;
    return tv;
  }
};

}  // namespace pgc
