"""Writes the golden fixtures under tests/golden/ from the strict CPU oracle.

Run from the repository root:  python tests/golden/make_golden.py
Each fixture holds the layered h/rx/cx circuit's instruction bytes (which pin the
generator, including the float angle bits), a seeded random initial state and the
oracle's final states from |0...0> and from that random state. From |0...0> the
layered circuit only multiplies |+>^N by a global phase (H^N|0> is an eigenvector
of every RX and CX), so the random-start state is the one that pins the kernels.
The tests cross-check every fixture against the numpy dense brute force (N <= 10)
and against the GPU.
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402

CASES = [  # (num_qubits, depth, seed)
    (4, 3, 11), (4, 10, 12), (8, 5, 21), (8, 10, 22), (12, 4, 31), (12, 10, 1001),
]


def main():
    oracle.build()
    index = []
    for n, d, seed in CASES:
        ins = oracle.gen_layered(n, d, seed)
        psi, _ = oracle.execute(n, ins)
        rng = np.random.default_rng(seed)
        init = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
        init /= np.linalg.norm(init)
        psi_r, _ = oracle.execute(n, ins, amps=init)
        name = f"layered_n{n}_d{d}_s{seed}.npz"
        np.savez_compressed(os.path.join(HERE, name), instrs=ins.view(np.uint8).reshape(-1, 16), state=psi,
                            init_random=init, state_random=psi_r)
        index.append({"file": name, "num_qubits": n, "depth": d, "seed": seed, "gates": int(len(ins)),
                      "generator": "mt19937_64(seed); theta=ldexp(rng()>>11,-53)*2pi as float; H all, "
                                   "then per layer RX all + CX(q,q+1) for q=l%2,l%2+2,..",
                      "oracle": "oracle/_build/liboracle.so (g++ -O2 -ffp-contract=off)"})
    with open(os.path.join(HERE, "index.json"), "w") as f:
        json.dump(index, f, indent=1)
    print(f"wrote {len(index)} fixtures")


if __name__ == "__main__":
    main()
