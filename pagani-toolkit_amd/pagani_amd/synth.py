"""Seeded synthetic voxel time series for the construction path.

Latent-community model: K factor series f_c(t) ~ N(0,1), each voxel i belongs to
one community c(i) and loads on it with L_i ~ U(0.3, 0.9):

    x_i(t) = L_i * f_c(i)(t) + 0.8 * eps_i(t),    eps ~ N(0,1)

computed in FP64 and stored FP32, time-major ``[T][N]`` exactly as the
reference gathers BOLD[l*N+i] from a NIfTI volume
(U/CorMat/CUCorMat/CUCORMAT/CorMat.cpp:223-228).

Random numbers: SplitMix64 (the k-th output of a stream is a closed form of k,
so every stream is generated vectorised), 53-bit uniforms in (0,1) and
Box-Muller pairs.  Each component draws from its own stream, keyed by
(seed, tag).
"""
from __future__ import annotations

import numpy as np

_GAMMA = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)
_TAGMUL = 0xD1B54A32D192ED03
_MASK = (1 << 64) - 1

TAG_FACTORS, TAG_COMMUNITY, TAG_LOADING, TAG_NOISE, TAG_SOURCES = 1, 2, 3, 4, 5


def _mix(z: np.ndarray) -> np.ndarray:
    z = (z ^ (z >> np.uint64(30))) * _M1
    z = (z ^ (z >> np.uint64(27))) * _M2
    return z ^ (z >> np.uint64(31))


def _stream_state(seed: int, tag: int) -> np.uint64:
    s = np.array([(seed ^ ((tag * _TAGMUL) & _MASK)) & _MASK], dtype=np.uint64)
    with np.errstate(over="ignore"):
        return _mix(s)[0]


def splitmix64(seed: int, tag: int, count: int) -> np.ndarray:
    """Outputs 1..count of the SplitMix64 stream keyed by (seed, tag)."""
    st = _stream_state(seed, tag)
    k = np.arange(1, count + 1, dtype=np.uint64)
    with np.errstate(over="ignore"):
        return _mix(st + k * _GAMMA)


def uniforms(seed: int, tag: int, count: int) -> np.ndarray:
    """Doubles in the open interval (0, 1)."""
    z = splitmix64(seed, tag, count)
    return ((z >> np.uint64(11)).astype(np.float64) + 0.5) * (1.0 / 9007199254740992.0)


def normals(seed: int, tag: int, count: int) -> np.ndarray:
    """Box-Muller standard normals (FP64)."""
    m = (count + 1) // 2
    u = uniforms(seed, tag, 2 * m)
    r = np.sqrt(-2.0 * np.log(u[0::2]))
    th = 2.0 * np.pi * u[1::2]
    out = np.empty(2 * m)
    out[0::2] = r * np.cos(th)
    out[1::2] = r * np.sin(th)
    return out[:count]


def communities(seed: int, n: int, k: int) -> np.ndarray:
    """c(i) = perm[i] mod K for a seeded random permutation (balanced sizes)."""
    perm = np.argsort(uniforms(seed, TAG_COMMUNITY, n), kind="stable")
    return (perm % k).astype(np.int64)


def bold(seed: int, n: int, t: int, k: int, noise: float = 0.8,
         dtype=np.float32) -> np.ndarray:
    """Synthetic BOLD matrix, time-major [T][N]."""
    f = normals(seed, TAG_FACTORS, k * t).reshape(k, t)
    c = communities(seed, n, k)
    load = 0.3 + 0.6 * uniforms(seed, TAG_LOADING, n)
    out = np.empty((t, n), dtype=dtype)
    # chunk over voxels to bound the FP64 temporaries
    step = 1 << 15
    for a in range(0, n, step):
        b = min(n, a + step)
        eps = normals(seed * 1000003 + a, TAG_NOISE, (b - a) * t).reshape(b - a, t)
        x = load[a:b, None] * f[c[a:b]] + noise * eps
        out[:, a:b] = x.T.astype(dtype)
    return out


def sample_sources(n: int, count: int, seed: int = 4242) -> np.ndarray:
    """`count` distinct vertices: a prefix of a seeded random permutation."""
    perm = np.argsort(uniforms(seed, TAG_SOURCES, n), kind="stable")
    return np.ascontiguousarray(perm[:count].astype(np.int32))


TAG_COHORT = 6


def cohort_subject(s: int):
    """(seed, voxels, time points) of cohort subject s: seed 4000 + s,
    N ~ U[20000, 70000] and T ~ U[150, 300] from that seed's own stream."""
    seed = 4000 + s
    u = uniforms(seed, TAG_COHORT, 2)
    return seed, 20000 + int(u[0] * 50001), 150 + int(u[1] * 151)


# Workload definitions (voxels, time points, communities, seed).
WORKLOADS = {
    "plain_n5000_t200": dict(n=5000, t=200, k=10, seed=1),
    "corr_n58000_t240": dict(n=58000, t=240, k=90, seed=2),
    "weighted_n120000_t300": dict(n=120000, t=300, k=200, seed=5),
}
