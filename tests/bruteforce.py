"""Independent dense brute force: each gate as an explicit 2^N x 2^N numpy matrix.

Gate matrices are written from their textbook definitions and expanded onto the
full space by index arithmetic in numpy, independently of both the oracle's and
the GPU's bit-insertion loops. Usable for N <= 10.
"""
from __future__ import annotations

import numpy as np

SQ2 = 1.0 / np.sqrt(2.0)


def _f32(u: int) -> float:
    return float(np.array([u], np.uint32).view(np.float32)[0])


def gate_matrix(g, pool=None, meta=None):
    """Returns (matrix, qubits) with matrix index bit i <-> qubits[i]."""
    t = int(g["type"])
    q = [int(x) for x in g["qubits"]]
    a = int(g["flags_arity"]) & 0xF
    th = _f32(int(g["payload"]))
    if t == 0x00:
        return np.array([[1, 1], [1, -1]], complex) * SQ2, q[:1]
    if t == 0x01:
        return np.array([[0, 1], [1, 0]], complex), q[:1]
    if t == 0x02:
        return np.array([[0, -1j], [1j, 0]], complex), q[:1]
    if t == 0x03:
        return np.array([[1, 0], [0, -1]], complex), q[:1]
    if t == 0x04:
        return np.array([[1, 0], [0, 1j]], complex), q[:1]
    if t in (0x10, 0x11, 0x12):
        c, s = np.cos(th / 2.0), np.sin(th / 2.0)
        if t == 0x10:
            return np.array([[c, -1j * s], [-1j * s, c]]), q[:1]
        if t == 0x11:
            return np.array([[c, -s], [s, c]], complex), q[:1]
        return np.diag([np.exp(-0.5j * th), np.exp(0.5j * th)]), q[:1]
    if t in (0x20, 0x21, 0xE1, 0x30, 0x31):
        k = {0x20: 2, 0x21: 2, 0xE1: 2, 0x30: 3, 0x31: 3}[t]
        M = np.zeros((1 << k, 1 << k), complex)
        for col in range(1 << k):
            b = [(col >> i) & 1 for i in range(k)]
            if t == 0x20 and b[0]:
                b[1] ^= 1
            elif t in (0x21, 0xE1):
                b[0], b[1] = b[1], b[0]
            elif t == 0x30 and b[0] and b[1]:
                b[2] ^= 1
            elif t == 0x31 and b[0]:
                b[1], b[2] = b[2], b[1]
            row = sum(v << i for i, v in enumerate(b))
            M[row, col] = 1.0
        return M, q[:k]
    if t in (0x80, 0xE0):
        e = meta[int(g["payload"])]
        d = 1 << a
        M = np.asarray(pool[int(e["offset"]):int(e["offset"]) + d * d]).reshape(d, d)
        qs = q[:min(a, 4)] + ([int(g["pad"]) & 0xFF] if a == 5 else [])
        return M, qs
    raise ValueError(f"no brute-force matrix for type {t:#x}")


def expand(M: np.ndarray, qubits, n: int) -> np.ndarray:
    idx = np.arange(1 << n)
    sub = np.zeros_like(idx)
    clear = idx.copy()
    for i, q in enumerate(qubits):
        bit = (idx >> q) & 1
        sub |= bit << i
        clear &= ~(1 << q)
    return M[sub[:, None], sub[None, :]] * (clear[:, None] == clear[None, :])


def run(n: int, instrs, pool=None, meta=None, psi=None, gswap: str = "noop") -> np.ndarray:
    if psi is None:
        psi = np.zeros(1 << n, complex)
        psi[0] = 1.0
    psi = np.array(psi, complex)
    for g in instrs:
        t = int(g["type"])
        if t == 0xF0 or (t == 0xE1 and gswap == "noop"):
            continue
        if t == 0xF1:
            raise ValueError("brute force does not sample measurements")
        M, qs = gate_matrix(g, pool, meta)
        psi = expand(M, qs, n) @ psi
    return psi


def clustered_state(n: int, cluster: int, depth: int, cluster_instrs) -> np.ndarray:
    """The clustered benchmark circuit never entangles two clusters, so from |0...0> its
    state is the Kronecker product of each cluster's own state (low qubits = low bits:
    cluster 0 is the rightmost factor). cluster_instrs(size) gives the instructions of one
    cluster of `size` qubits relabelled to qubits 0..size-1; each factor is brute-forced."""
    out = np.ones(1, complex)
    for lo in range(0, n, cluster):
        size = min(cluster, n - lo)
        out = np.kron(run(size, cluster_instrs(size)), out)
    return out


def random_state(n: int, rng: np.random.Generator) -> np.ndarray:
    v = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
    return v / np.linalg.norm(v)


def random_unitary(k: int, rng: np.random.Generator) -> np.ndarray:
    d = 1 << k
    z = rng.normal(size=(d, d)) + 1j * rng.normal(size=(d, d))
    q, r = np.linalg.qr(z)
    return q * (np.diag(r) / np.abs(np.diag(r)))
