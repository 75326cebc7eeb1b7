"""Deterministic random streams used at set-up time.

* ``ChaCha8Stream`` restates rand_chacha 0.3 ``ChaCha8Rng::seed_from_u64``
  (the reference's velocity RNG, src/runner.rs:3207-3213; pinned by
  Cargo.lock): rand_core 0.6 ``seed_from_u64`` expands the u64 seed into a
  32-byte key with PCG32, the block function is ChaCha with 8 rounds, a
  64-bit block counter in words 12-13 and stream 0 in words 14-15, and
  ``next_u64`` joins two consecutive little-endian u32 words.  rand 0.8's
  ``gen::<f64>()`` is ``(u64 >> 11) * 2^-53``.
* ``maxwell_boltzmann`` restates ``generate_velocities``
  (src/runner.rs:3194-3283).

The block function is vectorised over blocks with numpy so that ~10^6
velocity components take milliseconds.
"""
from __future__ import annotations

import numpy as np

_MASK64 = (1 << 64) - 1


def pcg32_seed_expand(seed: int, n_words: int = 8):
    """rand_core 0.6 SeedableRng::seed_from_u64 key expansion."""
    mul = 6364136223846793005
    inc = 11634580027462260723
    state = seed & _MASK64
    words = []
    for _ in range(n_words):
        state = (state * mul + inc) & _MASK64
        xorshifted = (((state >> 18) ^ state) >> 27) & 0xFFFFFFFF
        rot = state >> 59
        x = ((xorshifted >> rot) | (xorshifted << ((32 - rot) & 31))) & 0xFFFFFFFF
        words.append(x)
    return words


def _rotl(x, r):
    return ((x << np.uint32(r)) | (x >> np.uint32(32 - r))).astype(np.uint32)


def chacha_blocks(key_words, counters, stream=0, rounds=8, nonce_words=None):
    """ChaCha block function for each 64-bit counter in ``counters``.

    Returns an array of shape (len(counters), 16) of u32 output words.
    ``nonce_words`` (4 words for slots 12..15) overrides counter/stream, for
    the RFC 7539 known-answer vector."""
    counters = np.asarray(counters, dtype=np.uint64)
    nb = counters.shape[0]
    st = np.zeros((16, nb), np.uint32)
    st[0] = 0x61707865
    st[1] = 0x3320646E
    st[2] = 0x79622D32
    st[3] = 0x6B206574
    for i in range(8):
        st[4 + i] = key_words[i]
    if nonce_words is not None:
        for i in range(4):
            st[12 + i] = nonce_words[i]
    else:
        st[12] = (counters & np.uint64(0xFFFFFFFF)).astype(np.uint32)
        st[13] = (counters >> np.uint64(32)).astype(np.uint32)
        st[14] = stream & 0xFFFFFFFF
        st[15] = (stream >> 32) & 0xFFFFFFFF
    x = [st[i].copy() for i in range(16)]

    def qr(a, b, c, d):
        x[a] = x[a] + x[b]; x[d] = _rotl(x[d] ^ x[a], 16)
        x[c] = x[c] + x[d]; x[b] = _rotl(x[b] ^ x[c], 12)
        x[a] = x[a] + x[b]; x[d] = _rotl(x[d] ^ x[a], 8)
        x[c] = x[c] + x[d]; x[b] = _rotl(x[b] ^ x[c], 7)

    with np.errstate(over="ignore"):
        for _ in range(rounds // 2):
            qr(0, 4, 8, 12); qr(1, 5, 9, 13); qr(2, 6, 10, 14); qr(3, 7, 11, 15)
            qr(0, 5, 10, 15); qr(1, 6, 11, 12); qr(2, 7, 8, 13); qr(3, 4, 9, 14)
        out = np.stack([x[i] + st[i] for i in range(16)], axis=1)
    return out.astype(np.uint32)


class ChaCha8Stream:
    """ChaCha8Rng::seed_from_u64(seed) as a u64 / f64 stream."""

    def __init__(self, seed: int):
        self.key = pcg32_seed_expand(seed)
        self.counter = 0

    def next_u64_array(self, count: int) -> np.ndarray:
        n_words = 2 * count
        n_blocks = (n_words + 15) // 16
        blocks = chacha_blocks(self.key, np.arange(self.counter, self.counter + n_blocks,
                                                   dtype=np.uint64))
        self.counter += n_blocks
        words = blocks.reshape(-1)[:n_words].astype(np.uint64)
        return words[0::2] | (words[1::2] << np.uint64(32))

    def next_f64_array(self, count: int) -> np.ndarray:
        u = self.next_u64_array(count)
        return (u >> np.uint64(11)).astype(np.float64) * (1.0 / (1 << 53))


def maxwell_boltzmann(masses, kt, seed, n_constraints=0):
    """Velocities as in src/runner.rs:3194-3283 (f32 results).

    masses: per-atom masses (atomic units); kt: k_B*T in Hartree."""
    m = np.asarray(masses, dtype=np.float64)
    n = m.shape[0]
    v = np.zeros((3, n), np.float32)
    if kt == 0.0 or n == 0:
        return v
    # NOTE: the ChaCha8 buffer is consumed strictly in u64 pairs, so the
    # stream here is the keystream read as consecutive u64 values.
    u = ChaCha8Stream(seed).next_f64_array(6 * n).reshape(n, 3, 2)
    u1 = 1.0 - u[:, :, 0]
    u2 = u[:, :, 1]
    z = np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * np.pi * u2)
    sigma = np.sqrt(kt / m)
    v = (z * sigma[:, None]).astype(np.float32).T.copy()
    # sums in index order, as the reference's iterator sums (a cumulative
    # sum adds left to right; np.sum would add pairwise)
    total = float(np.cumsum(m)[-1])
    if total > 0:
        for ax in range(3):
            p = float(np.cumsum(m * v[ax].astype(np.float64))[-1])
            v[ax] = (v[ax].astype(np.float64) - p / total).astype(np.float32)
    ndof = 3 * n - n_constraints - 3
    if ndof <= 0:
        return np.zeros((3, n), np.float32)
    w = v.astype(np.float64)
    ke = float(np.cumsum(0.5 * m * ((w[0] ** 2 + w[1] ** 2) + w[2] ** 2))[-1])
    if ke > 0:
        scale = np.sqrt(0.5 * ndof * kt / ke)
        v = (v.astype(np.float64) * scale).astype(np.float32)
    return v
