"""Counter-based random streams shared by the oracle and the HIP kernels.

The reference draws every random variable from NumPy PCG64 generators hung on
a ``RandomNode`` tree (``hermespy/core/random_node.py``); its per-drop streams
are only statistically reproducible (bits sources and noise realisations are
re-rooted, ``hermespy/modem/modem.py:537``, ``hermespy/simulation/rf/block.py:104``).
The batched drop path therefore uses Philox4x32-10 (Salmon et al., SC'11,
"Parallel random numbers: as easy as 1, 2, 3"), keyed by the campaign seed and
counted by (drop, stream, index), so that every random number of every drop can
be regenerated independently on the GPU and here.

Stream layout (must match ``hermespy_amd/csrc/hpx_rng.hpp``):

* ``ctr = (drop, stream, index, 0)``, ``key = (seed & 0xffffffff, seed >> 32)``.
* stream ``STREAM_REALIZATION``: standard normals of one channel realisation, in
  the order the reference consumes them (``hermespy/channel/consistent.py:410-432``,
  offsets registered at ``hermespy/channel/fading/fading.py:739-754``).
  Call ``j`` yields normals ``2j`` and ``2j+1`` by Box-Muller on two 53-bit uniforms.
* stream ``STREAM_BITS + t``: the data bits of transmit stream ``t``; word ``i`` is
  output ``i % 4`` of call ``i // 4``; bit ``k`` is bit ``k % 32`` (LSB first) of word ``k // 32``.
* stream ``STREAM_NOISE + r``: complex AWGN of receive antenna ``r``; sample ``n`` uses
  outputs (0,1) (``n`` even) or (2,3) (``n`` odd) of call ``n // 2``, Box-Muller on
  24-bit uniforms.
"""

from __future__ import annotations

import numpy as np

PHILOX_M0 = 0xD2511F53
PHILOX_M1 = 0xCD9E8D57
PHILOX_W0 = 0x9E3779B9
PHILOX_W1 = 0xBB67AE85

STREAM_REALIZATION = 0x000
STREAM_BITS = 0x100
STREAM_NOISE = 0x200

_MASK = np.uint64(0xFFFFFFFF)


def philox4x32(counter: np.ndarray, key: np.ndarray, rounds: int = 10) -> np.ndarray:
    """Philox4x32 block function.

    Args:
        counter: uint32 array of shape (..., 4).
        key: uint32 array broadcastable to (..., 2).

    Returns: uint32 array of shape (..., 4).
    """

    counter = np.asarray(counter, dtype=np.uint32)
    key = np.broadcast_to(np.asarray(key, dtype=np.uint32), counter.shape[:-1] + (2,))
    c0, c1, c2, c3 = (counter[..., i].astype(np.uint64) for i in range(4))
    k0 = key[..., 0].astype(np.uint64)
    k1 = key[..., 1].astype(np.uint64)
    m0 = np.uint64(PHILOX_M0)
    m1 = np.uint64(PHILOX_M1)
    for r in range(rounds):
        if r > 0:
            k0 = (k0 + np.uint64(PHILOX_W0)) & _MASK
            k1 = (k1 + np.uint64(PHILOX_W1)) & _MASK
        p0 = m0 * c0
        p1 = m1 * c2
        c0, c1, c2, c3 = ((p1 >> np.uint64(32)) ^ c1 ^ k0, p1 & _MASK, (p0 >> np.uint64(32)) ^ c3 ^ k1, p0 & _MASK)
    return np.stack([c0, c1, c2, c3], axis=-1).astype(np.uint32)


def seed_key(seed: int) -> np.ndarray:
    seed = int(seed) & 0xFFFFFFFFFFFFFFFF
    return np.array([seed & 0xFFFFFFFF, seed >> 32], dtype=np.uint32)


def stream_words(seed: int, drop: int, stream: int, num_calls: int) -> np.ndarray:
    """Outputs of calls 0..num_calls-1 of one (drop, stream) pair, shape (num_calls, 4)."""

    ctr = np.zeros((num_calls, 4), dtype=np.uint32)
    ctr[:, 0] = np.uint32(drop)
    ctr[:, 1] = np.uint32(stream)
    ctr[:, 2] = np.arange(num_calls, dtype=np.uint32)
    return philox4x32(ctr, seed_key(seed))


def uniform53(hi: np.ndarray, lo: np.ndarray) -> np.ndarray:
    """Uniform on (0, 1) from two 32-bit words with 53 bits of resolution."""

    a = (hi.astype(np.uint64) >> np.uint64(5)).astype(np.float64)
    b = (lo.astype(np.uint64) >> np.uint64(6)).astype(np.float64)
    return (a * 67108864.0 + b + 0.5) * (1.0 / 9007199254740992.0)


def realization_normals(seed: int, drop: int, count: int) -> np.ndarray:
    """Standard normals of one drop's channel realisation (float64)."""

    calls = (count + 1) // 2
    w = stream_words(seed, drop, STREAM_REALIZATION, calls)
    u1 = uniform53(w[:, 0], w[:, 1])
    u2 = uniform53(w[:, 2], w[:, 3])
    rho = np.sqrt(-2.0 * np.log(u1))
    z = np.empty(2 * calls, dtype=np.float64)
    z[0::2] = rho * np.cos(2.0 * np.pi * u2)
    z[1::2] = rho * np.sin(2.0 * np.pi * u2)
    return z[:count]


def stream_bits(seed: int, drop: int, tx_stream: int, num_bits: int) -> np.ndarray:
    """Data bits (uint8 0/1) of one transmit stream of one drop."""

    num_words = (num_bits + 31) // 32
    calls = (num_words + 3) // 4
    words = stream_words(seed, drop, STREAM_BITS + tx_stream, calls).reshape(-1)[:num_words]
    bits = (words[:, None] >> np.arange(32, dtype=np.uint32)[None, :]) & np.uint32(1)
    return bits.reshape(-1)[:num_bits].astype(np.uint8)


def noise_normals(seed: int, drop: int, rx: int, num_samples: int) -> np.ndarray:
    """Unit-variance-per-component complex normals for receive antenna ``rx`` (complex128)."""

    calls = (num_samples + 1) // 2
    w = stream_words(seed, drop, STREAM_NOISE + rx, calls).reshape(-1, 2, 2)  # [call, pair, word]
    w = w.reshape(-1, 2)[:num_samples]
    u1 = ((w[:, 0] >> np.uint32(8)).astype(np.float64) + 0.5) * 2.0**-24
    u2 = (w[:, 1] >> np.uint32(8)).astype(np.float64) * 2.0**-24
    rho = np.sqrt(-2.0 * np.log(u1))
    return rho * np.cos(2.0 * np.pi * u2) + 1j * rho * np.sin(2.0 * np.pi * u2)
