"""CPU model of the symmetric blocked Hessian layout of k_minimize_nk (placement
MMFFHIP_HESS_GLOBAL_SYM) and of its H-times-vector pass.

The kernel stores the inverse Hessian as 4x4 blocks, one per unordered pair of row quads
{L, M}: block (s, L) pairs quad L with quad M = (L + s) mod nq, for the cyclic distances
s = 0 .. nq/2 (for even nq the pairs at distance nq/2 are stored once, by L < nq/2). Row r of
block (s, L) is float4 number (4 s + r) nq + L of the slot. Wave w of NW takes the distances
s = w, w + NW, ...; its lanes walk (distance, L) items in one of three mappings, and each item
adds its block's two products (B gM into rows 4L.., B^T gL into rows 4M..) to the wave's
private LDS accumulators (a float4 per quad) by a read, add and write. The tests below model
that walk: every block is visited once, no two lanes of one instruction update the same accumulator quad (so
a wave's read-add-write updates land in program order and the result is deterministic), the
slot and accumulator bounds hold, and the pass reproduces H g.
"""
from __future__ import annotations

import numpy as np
import pytest


def acc_floats_per_wave(max_n: int) -> int:
    """LDS floats reserved per wave for the accumulators (csrc sym_acc_floats)."""
    return max(256, max_n)


def slot_floats(max_n: int) -> int:
    nq = max_n // 4
    return 16 * (nq // 2 + 1) * nq


def wave_items(nq: int, w: int, NW: int):
    """Per step of wave w: a list of (lane, s, L, copy) for the lanes that process a block."""
    D = nq // 2 + 1
    Jw = (D - 1 - w) // NW + 1 if w < D else 0
    lanes = []
    for lane in range(64):
        if nq <= 64:
            G = 64 // nq
            copy = lane // nq
            L = lane - copy * nq
            j = copy if copy < G else Jw
        else:
            copy, j, L = 0, 0, lane
        lanes.append([j, L, copy])
    if nq <= 64:
        steps = (Jw + 64 // nq - 1) // (64 // nq)
    elif nq < 64 + NW:
        steps = 2 * Jw
    else:
        steps = (Jw * nq + 63) >> 6
    out = []
    for _ in range(steps):
        step = []
        for lane in range(64):
            j, L, copy = lanes[lane]
            if j < Jw and L < nq:
                s = w + j * NW
                if not (nq % 2 == 0 and s == D - 1 and s != 0 and L >= nq // 2):
                    step.append((lane, s, L, copy))
            # advance
            if nq <= 64:
                lanes[lane][0] = j + 64 // nq
            elif nq < 64 + NW:
                if L < 64:
                    lanes[lane][1] = L + 64
                else:
                    lanes[lane][1] = L - 64
                    lanes[lane][0] = j + 1
            else:
                L += 64
                if L >= nq:
                    L -= nq
                    j += 1
                lanes[lane][0], lanes[lane][1] = j, L
        out.append(step)
    return out


def copies(nq: int) -> int:
    return 64 // nq if nq <= 64 else 1


@pytest.mark.parametrize("NW", [1, 2, 4, 8])
def test_every_block_once_and_no_collisions(NW):
    for nq in list(range(1, 130)):
        n4 = 4 * nq
        seen = {}
        for w in range(NW):
            for step in wave_items(nq, w, NW):
                first, second = set(), set()
                for lane, s, L, copy in step:
                    assert copy < copies(nq) and copy * n4 + n4 <= acc_floats_per_wave(n4)
                    M = (L + s) % nq
                    key = (min(L, M), max(L, M))
                    assert key not in seen, (nq, NW, key)
                    seen[key] = (s, L)
                    assert (copy, L) not in first, (nq, NW, "first part collision")
                    first.add((copy, L))
                    if s != 0:
                        assert (copy, M) not in second, (nq, NW, "second part collision")
                        second.add((copy, M))
                    # float4 index of the block's rows inside the slot
                    assert (4 * s + 3) * nq + L < slot_floats(n4) // 4
        assert len(seen) == nq * (nq + 1) // 2, (nq, NW)


def _store(H: np.ndarray, nq: int) -> np.ndarray:
    """The slot of a symmetric matrix H (size <= 4 nq, zero padded) in the kernel's layout."""
    n = H.shape[0]
    D = nq // 2 + 1
    P = np.zeros((4 * nq, 4 * nq))
    P[:n, :n] = H
    slot = np.zeros((4 * D * nq, 4))
    for s in range(D):
        for r in range(4):
            for L in range(nq):
                M = (L + s) % nq
                slot[(4 * s + r) * nq + L] = P[4 * L + r, 4 * M:4 * M + 4]
    return slot


def _pass(slot: np.ndarray, g: np.ndarray, nq: int, NW: int) -> np.ndarray:
    """The kernel's H g: per-wave, per-copy accumulators (quad q at 4 q), summed in order."""
    n4 = 4 * nq
    gp = np.zeros(n4)
    gp[:len(g)] = g
    acc = np.zeros((NW, acc_floats_per_wave(n4)), np.float32)
    for w in range(NW):
        for step in wave_items(nq, w, NW):
            for lane, s, L, copy in step:
                M = (L + s) % nq
                B = np.stack([slot[(4 * s + r) * nq + L] for r in range(4)])
                o1 = B @ gp[4 * M:4 * M + 4]
                acc[w, copy * n4 + 4 * L:copy * n4 + 4 * L + 4] += o1
                if s:
                    o2 = B.T @ gp[4 * L:4 * L + 4]
                    acc[w, copy * n4 + 4 * M:copy * n4 + 4 * M + 4] += o2
    y = np.zeros(n4)
    for t in range(n4):
        for w in range(NW):
            for k in range(copies(nq)):
                y[t] += acc[w, k * n4 + t]
    return y[:len(g)]


@pytest.mark.parametrize("n,NW", [(9, 1), (30, 1), (96, 1), (99, 2), (150, 2), (192, 2), (258, 4),
                                  (270, 4), (300, 4), (360, 4), (363, 8), (471, 8)])
def test_pass_reproduces_h_times_g(n, NW):
    rng = np.random.default_rng(n)
    A = rng.normal(size=(n, n))
    H = A + A.T
    g = rng.normal(size=n)
    nq = (n + 3) // 4
    y = _pass(_store(H, nq), g, nq, NW)
    np.testing.assert_allclose(y, H @ g, rtol=1e-4, atol=1e-3 * np.abs(H @ g).max())
