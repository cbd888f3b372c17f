"""Seeded tiny ring instances on which the Lawler branch and bound branches.

(seed, |VX|, |VY|, max_depth) of synthetic.tiny_pair(ring=True). Found by a
seed scan: each explores at least three branches under one candidate pick
rule, and together they reach every branch-and-bound event the oracle counts
(ignored pushes, valid branches popped as the answer, the lambda_UB stop, an
empty queue, single-vertex drops, stale branches, sweep-cap calls, walks that
close on non-negative cycles). The last three are the instances on which the
GPU pick rule ends above the brute-force optimum while the CPU rule finds it.
"""
BRANCHING = [(717, 6, 6, 1), (1142, 5, 6, 1), (1736, 5, 6, 2), (1140, 6, 6, 1), (135, 6, 4, 2), (231, 6, 6, 1),
             (625, 5, 5, 1), (803, 5, 5, 1), (425, 6, 4, 2), (484, 6, 4, 2), (898, 6, 4, 2), (1241, 6, 4, 2),
             (3, 6, 4, 3), (1237, 6, 4, 3), (1268, 6, 4, 3), (898, 6, 4, 3), (83, 6, 4, 3), (90, 6, 4, 3),
             (1769, 6, 4, 1), (1513, 6, 4, 1), (1945, 6, 6, 1), (318, 6, 4, 1), (270, 6, 6, 1), (1461, 6, 6, 1),
             (1789, 6, 4, 1), (1610, 6, 4, 1), (12, 5, 4, 1), (83, 6, 4, 1), (83, 6, 4, 2)]
