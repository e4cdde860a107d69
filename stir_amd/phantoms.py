"""Seeded synthetic phantoms and Poisson data for the workloads.

Nothing here comes from files: every image and data set is generated from
a fixed seed so that the native path and the oracle receive identical
inputs.
"""
from __future__ import annotations

from typing import Callable, Tuple

import numpy as np

from .geometry import ProjDataInfoCylindrical, VoxelsOnCartesianGrid


def voxel_centres(img: VoxelsOnCartesianGrid):
    """(z, y, x) voxel-centre coordinates in mm, z relative to the image centre."""
    z = (np.arange(img.nz) - (img.nz - 1) / 2.0) * img.vox_z
    y = (np.arange(img.ny) + img.min_y) * img.vox_y
    x = (np.arange(img.nx) + img.min_x) * img.vox_x
    return np.meshgrid(z, y, x, indexing="ij")


def config0_phantom(img: VoxelsOnCartesianGrid, seed: int = 0) -> np.ndarray:
    """Uniform cylinder (radius 100 mm, value 1) plus 3 seeded spheres (radius 10 mm, value 4)."""
    rng = np.random.default_rng(seed)
    Z, Y, X = voxel_centres(img)
    out = np.where(X ** 2 + Y ** 2 <= 100.0 ** 2, 1.0, 0.0)
    zext = (img.nz - 1) / 2.0 * img.vox_z
    for _ in range(3):
        r = rng.uniform(0, 75.0)
        phi = rng.uniform(0, 2 * np.pi)
        cz = rng.uniform(-zext, zext)
        cx, cy = r * np.cos(phi), r * np.sin(phi)
        out = np.where((X - cx) ** 2 + (Y - cy) ** 2 + (Z - cz) ** 2 <= 10.0 ** 2, 4.0, out)
    return out.astype(np.float32)


def config1_phantom(img: VoxelsOnCartesianGrid, seed: int = 1) -> np.ndarray:
    """Ellipsoidal torso (1.0) with a cold lung region (0.2) and 10 hot spheres (3-8x)."""
    rng = np.random.default_rng(seed)
    Z, Y, X = voxel_centres(img)
    zext = (img.nz - 1) / 2.0 * img.vox_z
    a, b, c = 170.0, 120.0, zext * 1.05
    torso = (X / a) ** 2 + (Y / b) ** 2 + (Z / c) ** 2 <= 1.0
    out = np.where(torso, 1.0, 0.0)
    for sx in (-1, 1):
        lung = ((X - sx * 70.0) / 50.0) ** 2 + ((Y + 10.0) / 70.0) ** 2 + ((Z - 0.25 * zext) / (0.6 * zext)) ** 2 <= 1.0
        out = np.where(lung & torso, 0.2, out)
    for _ in range(10):
        rad = rng.uniform(6.0, 15.0)
        val = rng.uniform(3.0, 8.0)
        r = rng.uniform(0, 100.0)
        phi = rng.uniform(0, 2 * np.pi)
        cx, cy = 1.2 * r * np.cos(phi), 0.8 * r * np.sin(phi)
        cz = rng.uniform(-0.8 * zext, 0.8 * zext)
        out = np.where((X - cx) ** 2 + (Y - cy) ** 2 + (Z - cz) ** 2 <= rad ** 2, val, out)
    return out.astype(np.float32)


def poisson_data(forward: np.ndarray, mean_counts: float, background: float, seed: int) -> Tuple[np.ndarray, float]:
    """y ~ Poisson(scale * forward + background) with scale such that the
    scaled projection has mean `mean_counts` per bin. Returns (y, scale)."""
    rng = np.random.default_rng(seed)
    fwd = np.asarray(forward, dtype=np.float64)
    m = fwd.mean()
    scale = mean_counts / m if m > 0 else 1.0
    y = rng.poisson(scale * fwd + background).astype(np.float32)
    return y, float(scale)


def config0_data(pdi: ProjDataInfoCylindrical, img: VoxelsOnCartesianGrid,
                 forward_project: Callable[[np.ndarray], np.ndarray], seed: int = 0):
    """Phantom, data (Poisson, mean 5 counts/bin + background 0.1) and the background."""
    x = config0_phantom(img, seed)
    y, scale = poisson_data(forward_project(x), 5.0, 0.1, seed + 100)
    return x * np.float32(scale), y, 0.1


def config1_data(pdi: ProjDataInfoCylindrical, img: VoxelsOnCartesianGrid,
                 forward_project: Callable[[np.ndarray], np.ndarray], seed: int = 1):
    """Phantom, data (Poisson, mean 2 counts/bin + background 0.05) and the background."""
    x = config1_phantom(img, seed)
    y, scale = poisson_data(forward_project(x), 2.0, 0.05, seed + 100)
    return x * np.float32(scale), y, 0.05
