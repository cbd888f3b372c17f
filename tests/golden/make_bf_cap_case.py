"""Writes tests/golden/bf_cap_case.npz: a perturbed pair whose Lawler solve
runs every Bellman-Ford call of some branches to the numNodes-1 sweep cap
(26363 sweeps in total), with the oracle's answer.

Input: pair 111 of synthetic.make_uneven_pairs(256, seed=0) (icosphere s=3,
|VY|=41, max_depth 1) and one of its Gaussian feat_diff perturbations
(sigma 0.05, float32), given as a .npy file of float64 (exact float32 values)
on the command line. Output: the float32 feat_diff and the oracle's cycle
(edge ids), fp32 objective and counters (bf calls, sweeps, lawler
iterations, branches) of the lawlergpu rules.

    python tests/golden/make_bf_cap_case.py <feat_diff.npy>
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path[:0] = [os.path.join(HERE, "..", "..", "product-graph-cycles_amd"), os.path.join(HERE, "..", "..", "oracle")]

import pgc_oracle as O  # noqa: E402
from cyclic_product_graphs import synthetic as S  # noqa: E402


def main(path):
    p, K = S.make_uneven_pairs(256, seed=0)[111]
    fd = np.load(path)
    fd32 = fd.astype(np.float32)
    assert np.array_equal(fd32.astype(np.float64), fd), "feat_diff must hold float32 values"
    rows, src, trgt = O.product_space(p["EX"], p["EY"])
    en = O.energy(p["VX"], p["VY"], rows, fd32.astype(np.float64))
    e, obj, passes, stats, q = O.ctr_solve(src, trgt, en, len(p["VX"]), K, gpu_pick=True)
    np.savez_compressed(os.path.join(HERE, "bf_cap_case.npz"), feat_diff=fd32, pair_index=111, max_depth=K,
                        edges=np.asarray(e, np.int32), objective=np.float32(obj), passes=bool(passes),
                        stats=np.asarray(stats[:4], np.int64))
    print("edges", len(e), "objective", obj, "stats", stats)


if __name__ == "__main__":
    main(sys.argv[1])
