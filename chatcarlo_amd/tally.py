"""Voxel tallies and batch statistics (drop-in for chatcarlo/tally.py).

The arrays live on the host exactly as in the reference (C-order (nx, ny, nz)
float64 kerma [keV] and H*(10) track integral [pSv cm^3], plus the batch
moments Q and the per-voxel hit count).  Scoring happens on the GPU: the
transport kernels walk every flight segment through the grid with an exact
DDA and FP64 atomics into per-batch buffers, and a batch-end kernel folds a
batch into the totals and moments (chatcarlo/tally.py:59-94).  The
reference's stratified sub-step estimator has the exact segment-voxel overlap
as its expectation (chatcarlo/tally.py:255-266); the DDA scores that overlap
directly.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

_MEV_TO_JOULE = 1.602176634e-13
_G_TO_KG = 1e-3


@dataclass
class VoxelGrid:
    """chatcarlo/tally.py:26-132."""
    origin_cm: np.ndarray
    shape: tuple
    voxel_size_cm: float
    track_uncertainty: bool = False
    kerma_keV: np.ndarray = field(init=False)
    h10_track_pSv_cm3: np.ndarray = field(init=False)
    kerma_sum2: np.ndarray = field(init=False, default=None)
    h10_sum2: np.ndarray = field(init=False, default=None)
    n_batches_hit: np.ndarray = field(init=False, default=None)
    n_batches: int = field(init=False, default=0)
    n_histories: int = field(init=False, default=0)
    _kerma_prev: np.ndarray = field(init=False, default=None, repr=False)
    _h10_prev: np.ndarray = field(init=False, default=None, repr=False)

    def __post_init__(self):
        self.origin_cm = np.asarray(self.origin_cm, dtype=float)
        self.shape = tuple(int(x) for x in self.shape)
        self.kerma_keV = np.zeros(self.shape, dtype=float)
        self.h10_track_pSv_cm3 = np.zeros(self.shape, dtype=float)
        if self.track_uncertainty:
            self.kerma_sum2 = np.zeros(self.shape, dtype=float)
            self.h10_sum2 = np.zeros(self.shape, dtype=float)
            self.n_batches_hit = np.zeros(self.shape, dtype=np.int32)

    @classmethod
    def from_bbox(cls, bbox_min, bbox_max, resolution_cm: float, track_uncertainty: bool = False):
        """chatcarlo/tally.py:51-57: shape = max(1, ceil(extent / resolution))."""
        extent = np.asarray(bbox_max, dtype=float) - np.asarray(bbox_min, dtype=float)
        shape = tuple(max(1, int(np.ceil(x / resolution_cm))) for x in extent)
        return cls(origin_cm=np.asarray(bbox_min, dtype=float), shape=shape,
                   voxel_size_cm=resolution_cm, track_uncertainty=track_uncertainty)

    def end_batch(self, n_histories_in_batch: int) -> None:
        """Snapshot-difference batch moments on the host (chatcarlo/tally.py:59-94).

        run_transport folds batches on the GPU instead; this host form serves
        callers that drive transport_photons themselves.
        """
        if not self.track_uncertainty or n_histories_in_batch <= 0:
            return
        if self._kerma_prev is None:
            self._kerma_prev = np.zeros(self.shape, dtype=float)
            self._h10_prev = np.zeros(self.shape, dtype=float)
        dk = self.kerma_keV - self._kerma_prev
        dh = self.h10_track_pSv_cm3 - self._h10_prev
        self.kerma_sum2 += dk ** 2 / n_histories_in_batch
        self.h10_sum2 += dh ** 2 / n_histories_in_batch
        self.n_batches_hit += ((dk != 0) | (dh != 0)).astype(np.int32)
        np.copyto(self._kerma_prev, self.kerma_keV)
        np.copyto(self._h10_prev, self.h10_track_pSv_cm3)
        self.n_batches += 1
        self.n_histories += n_histories_in_batch

    def kerma_relative_error(self) -> np.ndarray:
        return relative_error(self.kerma_keV, self.kerma_sum2, self.n_batches, self.n_histories)

    def h10_relative_error(self) -> np.ndarray:
        return relative_error(self.h10_track_pSv_cm3, self.h10_sum2, self.n_batches, self.n_histories)

    def voxel_index(self, points: np.ndarray):
        idx = np.floor((points - self.origin_cm) / self.voxel_size_cm).astype(int)
        valid = np.all((idx >= 0) & (idx < np.array(self.shape)), axis=1)
        return idx, valid

    def voxel_centers(self) -> np.ndarray:
        nx, ny, nz = self.shape
        xs = self.origin_cm[0] + (np.arange(nx) + 0.5) * self.voxel_size_cm
        ys = self.origin_cm[1] + (np.arange(ny) + 0.5) * self.voxel_size_cm
        zs = self.origin_cm[2] + (np.arange(nz) + 0.5) * self.voxel_size_cm
        gx, gy, gz = np.meshgrid(xs, ys, zs, indexing="ij")
        return np.stack([gx.ravel(), gy.ravel(), gz.ravel()], axis=1)

    def voxel_volume_cm3(self) -> float:
        return self.voxel_size_cm ** 3

    def total_kerma_MeV(self) -> float:
        return float(self.kerma_keV.sum()) / 1000.0

    def dose_map_Gy(self, density_g_cm3: np.ndarray) -> np.ndarray:
        mass_kg = density_g_cm3 * self.voxel_volume_cm3() * _G_TO_KG
        energy_J = self.kerma_keV * 1e-3 * _MEV_TO_JOULE
        return np.divide(energy_J, mass_kg, out=np.zeros_like(energy_J), where=mass_kg > 0)

    def h10_map_pSv(self) -> np.ndarray:
        return self.h10_track_pSv_cm3 / self.voxel_volume_cm3()


def _batch_variance(T, Q, n_batches: int, n_histories: int):
    """sigma^2 = (Q - T^2/N) / (M - 1), clamped at 0 (chatcarlo/tally.py:135-150)."""
    T = np.asarray(T, dtype=float)
    Q = np.asarray(Q, dtype=float)
    if n_batches < 2 or n_histories <= 0:
        return np.full_like(T, np.nan)
    return np.clip((Q - T ** 2 / n_histories) / (n_batches - 1), 0.0, None)


def standard_error(T, Q, n_batches: int, n_histories: int):
    var = _batch_variance(T, Q, n_batches, n_histories)
    if n_histories <= 0:
        return np.full_like(var, np.nan)
    return np.sqrt(var / n_histories)


def relative_error(T, Q, n_batches: int, n_histories: int):
    T = np.asarray(T, dtype=float)
    sem = standard_error(T, Q, n_batches, n_histories)
    if n_histories <= 0:
        return np.full_like(T, np.nan)
    mean = T / n_histories
    with np.errstate(divide="ignore", invalid="ignore"):
        return np.where(mean != 0, sem / mean, np.nan)


def combine_moments(moments_a: tuple, moments_b: tuple) -> tuple:
    t_a, q_a, m_a, n_a = moments_a
    t_b, q_b, m_b, n_b = moments_b
    return (t_a + t_b, q_a + q_b, m_a + m_b, n_a + n_b)


@dataclass
class ScalarMoments:
    """Per-material batch moments (chatcarlo/tally.py:191-245)."""
    sums: dict = field(default_factory=dict)
    sum2s: dict = field(default_factory=dict)
    n_batches: int = 0
    n_histories: int = 0

    def add_batch(self, batch_values: dict, n_histories_in_batch: int) -> None:
        if n_histories_in_batch <= 0:
            return
        for name in set(self.sums) | set(batch_values):
            s_b = batch_values.get(name, 0.0)
            self.sums[name] = self.sums.get(name, 0.0) + s_b
            self.sum2s[name] = self.sum2s.get(name, 0.0) + s_b ** 2 / n_histories_in_batch
        self.n_batches += 1
        self.n_histories += n_histories_in_batch

    def relative_errors(self) -> dict:
        return {k: float(relative_error(self.sums[k], self.sum2s[k], self.n_batches, self.n_histories))
                for k in self.sums}

    def standard_errors(self) -> dict:
        return {k: float(standard_error(self.sums[k], self.sum2s[k], self.n_batches, self.n_histories))
                for k in self.sums}

    def as_moments(self) -> dict:
        return {k: (self.sums[k], self.sum2s[k], self.n_batches, self.n_histories) for k in self.sums}

    def merge_from(self, other_moments: dict, other_n_batches: int, other_n_histories: int) -> None:
        for k in set(self.sums) | set(other_moments):
            t, q = other_moments.get(k, (0.0, 0.0))[:2]
            self.sums[k] = self.sums.get(k, 0.0) + t
            self.sum2s[k] = self.sum2s.get(k, 0.0) + q
        self.n_batches += other_n_batches
        self.n_histories += other_n_histories
