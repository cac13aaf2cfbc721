"""Seeded workload generators (synthetic circuits of the benchmark configs).

Every generator returns a list of (U, quil_qubits) gates, U in the Quil
matrix convention.  Randomness comes from splitmix64 so the same seed gives
the same circuit on every machine; doubles are (x >> 11) * 2^-53.

* ``layered``   -- L layers of RX(a) q, RZ(b) q on every qubit (a, b uniform in
                   [0, 2pi)), then CNOT(i, i+1) for even i on even layers and odd
                   i on odd layers.  The structure of the reference's
                   benchmark circuits (app/src/benchmark-programs.lisp:105-126)
                   with deterministic brickwork.
* ``qft``       -- the QFT in the gate order of app/src/benchmark-programs.lisp:18-73
                   (H, CPHASE(pi/2^k) ladders, final bit-reversal SWAPs).
* ``haar_cz``   -- depth-D layers of a Haar-random U3 per qubit (Ginibre QR with
                   the phase fix) and CZ on a seeded random perfect matching.
* ``qaoa_maxcut`` -- H on all qubits, then p layers of RZZ(gamma) per edge and
                   RX(beta) per qubit (examples/qaoa.lisp:7-75).
* ``random_mix``-- 60% Haar 1q U3 / 40% CNOT-or-CZ on random pairs.
"""
from __future__ import annotations

import math
from typing import List, Sequence, Tuple

import numpy as np

from . import gates as G

MASK64 = (1 << 64) - 1


class SplitMix64:
    def __init__(self, seed: int):
        self.x = seed & MASK64

    def next_u64(self) -> int:
        self.x = (self.x + 0x9E3779B97F4A7C15) & MASK64
        z = self.x
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
        return z ^ (z >> 31)

    def uniform(self) -> float:
        return (self.next_u64() >> 11) * (1.0 / (1 << 53))

    def below(self, n: int) -> int:
        return int(self.uniform() * n)

    def normal(self) -> float:
        # Box-Muller on two uniforms (u1 in (0, 1])
        u1 = 1.0 - self.uniform()
        u2 = self.uniform()
        return math.sqrt(-2.0 * math.log(u1)) * math.cos(2 * math.pi * u2)


def uniforms(seed: int, n: int) -> np.ndarray:
    """n doubles in [0,1) from splitmix64 (vectorised)."""
    x = np.uint64(seed & MASK64) + np.arange(1, n + 1, dtype=np.uint64) * np.uint64(0x9E3779B97F4A7C15)
    with np.errstate(over="ignore"):
        z = x.copy()
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return (z >> np.uint64(11)).astype(np.float64) * (1.0 / (1 << 53))


Gate = Tuple[np.ndarray, List[int]]


def layered(n: int, layers: int, seed: int = 1) -> List[Gate]:
    rng = SplitMix64(seed)
    out: List[Gate] = []
    for L in range(layers):
        for q in range(n):
            a = 2 * math.pi * rng.uniform()
            b = 2 * math.pi * rng.uniform()
            out.append((G.gate_matrix("RX", [a]), [q]))
            out.append((G.gate_matrix("RZ", [b]), [q]))
        for i in range(L % 2, n - 1, 2):
            out.append((G.gate_matrix("CNOT"), [i, i + 1]))
    return out


def qft(n: int) -> List[Gate]:
    """QFT of app/src/benchmark-programs.lisp:18-73 on qubits 0..n-1."""
    def rec(qs: Sequence[int]) -> List[Gate]:
        q, rest = qs[0], list(qs[1:])
        if not rest:
            return [(G.gate_matrix("H"), [q])]
        nn = len(rest) + 1
        cr: List[Gate] = []
        for idx, qi in enumerate(rest):
            i = nn - 1 - idx
            angle = math.pi / (2 ** (nn - i))
            cr.insert(0, (G.gate_matrix("CPHASE", [angle]), [q, qi]))  # push
        return rec(rest) + cr + [(G.gate_matrix("H"), [q])]

    qs = list(range(n))
    out = rec(qs) if n else []
    if n >= 2:
        for i in range(n // 2):
            out.append((G.gate_matrix("SWAP"), [qs[i], qs[n - 1 - i]]))
    return out


def haar_u2(rng: SplitMix64) -> np.ndarray:
    """Haar-random 2x2 unitary: QR of a complex Ginibre matrix with the
    diagonal phase fix (Mezzadri 2007)."""
    Z = np.array([[complex(rng.normal(), rng.normal()) for _ in range(2)] for _ in range(2)]) / math.sqrt(2)
    Q, R = np.linalg.qr(Z)
    d = np.diag(R)
    return Q * (d / np.abs(d))


def haar_cz(n: int, depth: int, seed: int = 3) -> List[Gate]:
    rng = SplitMix64(seed)
    out: List[Gate] = []
    cz = G.gate_matrix("CZ")
    for _ in range(depth):
        for q in range(n):
            out.append((haar_u2(rng), [q]))
        perm = list(range(n))
        for i in range(n - 1, 0, -1):  # Fisher-Yates
            j = rng.below(i + 1)
            perm[i], perm[j] = perm[j], perm[i]
        for i in range(0, n - 1, 2):
            out.append((cz, [perm[i], perm[i + 1]]))
    return out


def random_regular3(n: int, rng: SplitMix64, tries: int = 1000) -> List[Tuple[int, int]]:
    """Random 3-regular simple graph on n (even) vertices: pairing model with
    rejection of loops and multi-edges."""
    if n % 2 or n < 4:
        raise ValueError("3-regular graphs need an even n >= 4")
    for _ in range(tries):
        pts = [v for v in range(n) for _ in range(3)]
        for i in range(len(pts) - 1, 0, -1):
            j = rng.below(i + 1)
            pts[i], pts[j] = pts[j], pts[i]
        edges = set()
        ok = True
        for i in range(0, len(pts), 2):
            a, b = pts[i], pts[i + 1]
            if a == b or (min(a, b), max(a, b)) in edges:
                ok = False
                break
            edges.add((min(a, b), max(a, b)))
        if ok:
            return sorted(edges)
    raise RuntimeError("no simple 3-regular graph found")


def qaoa_maxcut(n: int, edges, gammas, betas) -> List[Gate]:
    out: List[Gate] = [(G.gate_matrix("H"), [q]) for q in range(n)]
    for g, b in zip(gammas, betas):
        for (a, c) in edges:
            out.append((G.gate_matrix("RZZ", [g]), [a, c]))
        for q in range(n):
            out.append((G.gate_matrix("RX", [b]), [q]))
    return out


def qaoa_instances(count: int, seed: int = 4, nmin: int = 10, nmax: int = 24):
    """(n, edges, gammas, betas) per instance; n uniform over the even values
    in [nmin, nmax] (3-regular graphs need even n)."""
    rng = SplitMix64(seed)
    evens = [x for x in range(nmin, nmax + 1) if x % 2 == 0]
    out = []
    for _ in range(count):
        n = evens[rng.below(len(evens))]
        edges = random_regular3(n, rng)
        p = 1 + rng.below(4)
        gam = [math.pi * rng.uniform() for _ in range(p)]
        bet = [0.5 * math.pi * rng.uniform() for _ in range(p)]
        out.append((n, edges, gam, bet))
    return out


def random_mix(n: int, depth: int, seed: int = 5) -> List[Gate]:
    """depth * n gates: 60% Haar U3 on a random qubit, 40% CNOT or CZ on a
    random ordered pair."""
    rng = SplitMix64(seed)
    out: List[Gate] = []
    cnot, cz = G.gate_matrix("CNOT"), G.gate_matrix("CZ")
    for _ in range(depth * n):
        if rng.uniform() < 0.6:
            out.append((haar_u2(rng), [rng.below(n)]))
        else:
            a = rng.below(n)
            b = rng.below(n - 1)
            b = b + 1 if b >= a else b
            out.append((cnot if rng.uniform() < 0.5 else cz, [a, b]))
    return out


def haar_unitary(k: int, rng: SplitMix64) -> np.ndarray:
    """Haar-random 2^k x 2^k unitary (Ginibre QR with the phase fix)."""
    d = 1 << k
    Z = np.array([[complex(rng.normal(), rng.normal()) for _ in range(d)] for _ in range(d)]) / math.sqrt(2)
    Q, R = np.linalg.qr(Z)
    r = np.diag(R)
    return Q * (r / np.abs(r))


def dense_blocks(n: int, depth: int, k: int, seed: int = 7) -> List[Gate]:
    """depth layers of a Haar 1-qubit gate on every qubit followed by
    floor(n / k) Haar k-qubit unitaries on a seeded random partition of the
    qubits (dense k-qubit blocks, as a fusing compiler would emit)."""
    rng = SplitMix64(seed)
    out: List[Gate] = []
    for _ in range(depth):
        for q in range(n):
            out.append((haar_u2(rng), [q]))
        perm = list(range(n))
        for i in range(n - 1, 0, -1):
            j = rng.below(i + 1)
            perm[i], perm[j] = perm[j], perm[i]
        for b in range(n // k):
            out.append((haar_unitary(k, rng), perm[b * k:(b + 1) * k]))
    return out


def diagonal_heavy(n: int, depth: int, seed: int = 8) -> List[Gate]:
    """QAOA-like: H on every qubit, then per layer ZZ phases (RZZ) on 2n
    random pairs and RZ, then RX, on every qubit: mostly diagonal gates."""
    rng = SplitMix64(seed)
    out: List[Gate] = [(G.gate_matrix("H"), [q]) for q in range(n)]
    for _ in range(depth):
        for _ in range(2 * n):
            a = rng.below(n)
            b = rng.below(n - 1)
            b = b + 1 if b >= a else b
            out.append((G.gate_matrix("RZZ", [math.pi * rng.uniform()]), [a, b]))
        for q in range(n):
            out.append((G.gate_matrix("RZ", [2 * math.pi * rng.uniform()]), [q]))
        for q in range(n):
            out.append((G.gate_matrix("RX", [math.pi * rng.uniform()]), [q]))
    return out


def mixed_workload(scale: str = "test", seed: int = 900):
    """Circuits on which the best plan differs: (name, n, dtype, gates).
    "test" keeps every state at <= 2^26 amplitudes; "bench" has the
    n = 28..30 members of the families and one small state."""
    if scale == "bench":
        spec = [("random_mix", 30, "f64", 20), ("random_mix", 30, "f32", 20), ("dense_k6", 28, "f64", 4),
                ("diagonal_heavy", 30, "f64", 3), ("layered_small", 18, "f64", 40)]
    else:
        spec = [("random_mix", 26, "f64", 30), ("random_mix", 24, "f32", 30), ("dense_k6", 24, "f64", 5),
                ("dense_k5", 26, "f32", 5), ("diagonal_heavy", 26, "f64", 5), ("layered_small", 16, "f64", 30),
                ("dense_k6", 18, "f32", 8)]
    out = []
    for j, (name, n, dtype, depth) in enumerate(spec):
        s = seed + j
        if name == "random_mix":
            gates = random_mix(n, depth, seed=s)
        elif name.startswith("dense_k"):
            gates = dense_blocks(n, depth, int(name[-1]), seed=s)
        elif name == "diagonal_heavy":
            gates = diagonal_heavy(n, depth, seed=s)
        else:
            gates = layered(n, depth, seed=s)
        out.append((f"{name}_n{n}_{dtype}", n, dtype, gates))
    return out


def amp_updates(gates: List[Gate], n: int) -> int:
    """Sum of changed(g): amplitudes each gate can modify."""
    return sum(G.changed_amplitudes(U, n) for U, _ in gates)
