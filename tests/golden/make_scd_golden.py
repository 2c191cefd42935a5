"""Writes tests/golden/scd_downsample_golden.npz (development container only).

Runs the reference's own downsample_projection_and_list
(sparse_main_utils.py:130-160) on small seeded inputs and stores inputs and
outputs.  The count and size rules of that function are pinned by properties
(tests/test_oracle_pins.py); what only the reference's call pins is the order
in which it draws its torch.randperm permutations from the global generator
(per rotation: the bin copies first, then the list rows).
Usage: python tests/golden/make_scd_golden.py
"""
import os
import sys

import numpy as np
import torch

HERE = os.path.dirname(os.path.abspath(__file__))
REF = "/root/reference"
CASES = [(20260331, 0.5), (7, 0.3), (11, 0.05)]


def inputs(seed):
    rng = np.random.default_rng(seed)
    proj = rng.poisson(2.5, (23, 4)).astype(np.float32)
    proj[3, 1] += 0.5           # a half count: Python's round() rounds it to even
    proj[:, 2] = 0.0            # an empty rotation: no permutation drawn for its bins
    lists = [rng.normal(size=(int(rng.integers(1, 30)), 4)).astype(np.float32) for _ in range(4)]
    lists[3] = lists[3][:1]     # a one-row list: max(int(n * ds), 1) keeps it
    return proj, lists


def main():
    sys.path.insert(0, REF)
    import sparse_main_utils as ref  # the reference's module (CPU run only)
    out = {}
    for k, (seed, ds) in enumerate(CASES):
        proj, lists = inputs(seed)
        torch.manual_seed(seed)
        p_out, l_out = ref.downsample_projection_and_list(torch.from_numpy(proj), [torch.from_numpy(l) for l in lists],
                                                          ds)
        out[f"c{k}_seed"] = np.array(seed)
        out[f"c{k}_ds"] = np.array(ds)
        out[f"c{k}_proj_in"] = proj
        out[f"c{k}_proj_out"] = p_out.numpy()
        for r, (li, lo) in enumerate(zip(lists, l_out)):
            out[f"c{k}_list{r}_in"] = li
            out[f"c{k}_list{r}_out"] = lo.numpy()
    out["n_cases"] = np.array(len(CASES))
    np.savez_compressed(os.path.join(HERE, "scd_downsample_golden.npz"), **out)
    print("wrote", len(CASES), "cases")


if __name__ == "__main__":
    main()
