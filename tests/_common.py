"""Shared helpers of the parity tests: seeded inputs and comparisons."""
from __future__ import annotations

import numpy as np

from softbox_amd import scenes


def assert_same(got: np.ndarray, want: np.ndarray, what: str) -> None:
    """Bit-exact for integers; exact value equality for floats (-0.0 == 0.0, NaN == NaN)."""
    got = np.asarray(got)
    want = np.asarray(want)
    assert got.shape == want.shape, f"{what}: shape {got.shape} != {want.shape}"
    if np.issubdtype(want.dtype, np.floating):
        bad = ~((got == want) | (np.isnan(got) & np.isnan(want)))
    else:
        bad = got != want
    nb = int(bad.sum())
    if nb:
        idx = np.flatnonzero(bad)[:8]
        raise AssertionError(f"{what}: {nb}/{got.size} differ; first at {idx.tolist()}: "
                             f"got {got.ravel()[idx].tolist()} want {want.ravel()[idx].tolist()}")


def max_rel(got: np.ndarray, want: np.ndarray, floor: float = 0.0) -> float:
    got = np.asarray(got, np.float64)
    want = np.asarray(want, np.float64)
    scale = np.maximum(np.abs(want), floor)
    with np.errstate(divide="ignore", invalid="ignore"):
        r = np.where(scale > 0, np.abs(got - want) / scale, np.abs(got - want))
    return float(r.max()) if r.size else 0.0


def rod_state(n: int, seed: int, forces: bool = True):
    """Free rods with random lab-frame forces/torques of realistic magnitude (for integrate tests)."""
    sc = scenes.free_rods(n=n, seed=seed, grid=False, periodic=True)
    fs = sc.fs
    g = np.random.default_rng(seed + 1000)
    if forces:
        fs.forceSum[:] = (g.standard_normal(3 * n) * 1e-12).astype(np.float32)
        fs.torqueSum[:] = (g.standard_normal(3 * n) * 1e-19).astype(np.float32)
        fs.randForce[:] = (g.standard_normal(3 * n) * 1e-12).astype(np.float32)
        fs.randTorque[:] = (g.standard_normal(3 * n) * 1e-19).astype(np.float32)
    fs.brownTransScale[:] = g.random(n).astype(np.float32) + 0.5
    fs.brownRotScale[:] = g.random(n).astype(np.float32) + 0.5
    return sc


def small_network(seed: int = 1, n_fil: int = 200, links: int = 500, periodic: bool = True, **kw):
    box = 20.0 * (n_fil / 20000.0) ** (1.0 / 3.0)
    return scenes.network(n_fil=n_fil, segs=10, box=box, n_links=links, seed=seed, periodic=periodic, **kw)
