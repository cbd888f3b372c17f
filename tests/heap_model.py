"""The reference's branch-and-bound queue restated in Python (test infrastructure).

An independent restatement of helper/minheap.cpp (constructor :16-31,
decrease :37-52, peakKey :77-82, pop :87-116, push :137-159, bubbleDown
:169-235, bubbleUp :239-262, switchEntries :266-277; peak helper/minheap.hpp:31-33),
step for step, so that the C++ heaps of the oracle (oracle/minheap_oracle.hpp)
and of the product's host driver (csrc/bnb_heap.h) can be checked against it
on operation sequences. Pop order among equal keys, the ignored pushes and
the holes left by pop all follow from these steps.
"""
INF = float("inf")


class RefMinHeap:
    def __init__(self, max_size):
        self.heap = [[-1.0, -1] for _ in range(max_size)]
        self.map = [-1] * max_size
        self.max_size = max_size
        self.size = 0
        self.min_empty = 0
        self.max_full = -1

    def peak(self):
        return INF if self.size == 0 else self.heap[0][0]

    def peak_key(self, key):
        if key < 0 or key >= self.max_size or self.map[key] == -1:
            return -1.0
        return self.heap[self.map[key]][0]

    def decrease(self, key, value):
        if key < 0 or key >= self.max_size:
            return False
        i = self.map[key]
        if i != -1 and value < self.heap[i][0]:
            self.heap[i][0] = value
            self._bubble_up(i)
            if self.heap[i][0] != value:
                return True
        return False

    def pop(self):
        if self.size == 0:
            return (-1.0, -1)
        minimum = (self.heap[0][0], self.heap[0][1])
        self.heap[0][0] = INF
        self._switch(0, self.max_full)
        self.heap[self.max_full] = [-1.0, -1]
        self.map[minimum[1]] = -1
        new_index = self._bubble_down(0)
        if self.max_full < self.min_empty:
            self.min_empty = self.max_full
        while self.max_full >= 0 and self.heap[self.max_full][1] == -1:
            self.max_full -= 1
        if new_index > self.max_full:
            self.max_full = new_index
        self.size -= 1
        return minimum

    def push(self, value, key):
        if key < 0 or key >= self.max_size or self.map[key] != -1 or self.size == self.max_size:
            return
        self.heap[self.min_empty] = [value, key]
        self.map[key] = self.min_empty
        self._bubble_up(self.min_empty)
        self.size += 1
        if self.min_empty == self.max_full + 1:
            self.max_full += 1
        if self.size == self.max_size:
            self.min_empty = self.size + 1
            return
        while self.heap[self.min_empty][1] > -1:
            self.min_empty += 1

    def _bubble_down(self, i):
        if i < 0 or i >= self.max_size:
            return -1
        h, n = self.heap, self.max_size
        left, right = 2 * i + 1, 2 * i + 2
        while (left < n or right < n) and (left != -1 or right != -1):
            lp = left != -1 and left < n and h[left][1] != -1 and h[left][0] <= h[i][0]
            rp = right != -1 and right < n and h[right][1] != -1 and h[right][0] <= h[i][0]
            go_left = go_right = False
            if lp and rp:
                if h[right][0] <= h[left][0]:
                    go_right = True
                else:
                    go_left = True
            elif lp:
                go_left = True
            elif rp:
                go_right = True
            if go_left or go_right:
                c = left if go_left else right
                h[c], h[i] = h[i], h[c]
                self.map[h[i][1]] = i
                self.map[h[c][1]] = c
                i = c
                left, right = 2 * i + 1, 2 * i + 2
            else:
                left = right = -1
        return i

    def _bubble_up(self, i):
        if i < 0 or i >= self.max_size:
            return -1
        h = self.heap
        parent = -1 if i <= 0 else (i - 1) // 2
        while parent != -1:
            if h[parent][0] >= h[i][0]:
                h[parent], h[i] = h[i], h[parent]
                self.map[h[i][1]] = i
                self.map[h[parent][1]] = parent
                i = parent
                parent = -1 if i <= 0 else (i - 1) // 2
            else:
                parent = -1
        return i

    def _switch(self, i, j):
        h = self.heap
        if i == j or i < 0 or j < 0 or i >= self.max_size or j >= self.max_size or h[i][1] == -1 or h[j][1] == -1:
            return
        h[i], h[j] = h[j], h[i]
        self.map[h[j][1]] = j
        self.map[h[i][1]] = i


# Operation codes shared with the C++ runners (tests/heap_harness.cpp and
# orc_heap_run): 0 push(value, key), 1 pop, 2 peak, 3 decrease(key, value),
# 4 peakKey(key). Every operation but push records one (value, key) output.
def run(cap, ops):
    h = RefMinHeap(cap)
    out = []
    for op, key, val in ops:
        if op == 0:
            h.push(val, key)
        elif op == 1:
            out.append(h.pop())
        elif op == 2:
            out.append((h.peak(), -2))
        elif op == 3:
            out.append((1.0 if h.decrease(key, val) else 0.0, -3))
        else:
            out.append((h.peak_key(key), -4))
    return out
