// Test harness (not product code): runs an operation sequence on a C++
// branch-and-bound heap and records its outputs, so that tests/test_heap.py can
// compare it with tests/heap_model.py. Built by the test with g++ against
// the header named by HEAP_HEADER (the product's csrc/bnb_heap.h, namespace
// pgc) or the oracle's minheap_oracle.hpp (namespace oracle), HEAP_NS.
#include HEAP_HEADER

extern "C" long heap_run(long cap, const int* op, const long* key, const double* val, long n, double* out_v,
                         long* out_k) {
  HEAP_NS::BnbHeap h(cap);
  long m = 0;
  for (long i = 0; i < n; ++i) {
    switch (op[i]) {
      case 0:
        h.push(val[i], key[i]);
        break;
      case 1: {
        const auto t = h.pop();
        out_v[m] = t.first;
        out_k[m++] = t.second;
        break;
      }
      case 2:
        out_v[m] = h.peak();
        out_k[m++] = -2;
        break;
      case 3:
        out_v[m] = h.decrease(key[i], val[i]) ? 1.0 : 0.0;
        out_k[m++] = -3;
        break;
      default:
        out_v[m] = h.key_value(key[i]);
        out_k[m++] = -4;
        break;
    }
  }
  return m;
}
