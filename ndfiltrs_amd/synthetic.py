"""Seeded synthetic astronomical images for the benchmark workloads.

Each generator returns a C-contiguous array drawn with NumPy's PCG64 from a
fixed seed, so every run (and the parity checks) see the same pixels.
"""

from __future__ import annotations

import numpy as np

BASE_SEED = 20261203


def _add_sources(image: np.ndarray, rng: np.random.Generator, count: int) -> None:
    """Circular Gaussians, sigma U(1.5, 4) px, peak log-U(50, 5000), uniform positions."""
    rows, cols = image.shape[-2:]
    sigmas = rng.uniform(1.5, 4.0, size=count)
    peaks = np.exp(rng.uniform(np.log(50.0), np.log(5000.0), size=count))
    ys = rng.uniform(0, rows, size=count)
    xs = rng.uniform(0, cols, size=count)
    for sigma, peak, y, x in zip(sigmas, peaks, ys, xs):
        radius = int(np.ceil(5 * sigma))
        y0, y1 = max(int(y) - radius, 0), min(int(y) + radius + 1, rows)
        x0, x1 = max(int(x) - radius, 0), min(int(x) + radius + 1, cols)
        yy, xx = np.mgrid[y0:y1, x0:x1]
        image[..., y0:y1, x0:x1] += peak * np.exp(-((yy - y) ** 2 + (xx - x) ** 2) / (2 * sigma**2))


def sky_small(shape=(2048, 2048), seed: int = BASE_SEED + 1) -> np.ndarray:
    """5x5 reflect workload: N(1000,10) sky, 200 sources, 1% NaN, float32."""
    rng = np.random.default_rng(seed)
    image = rng.normal(1000.0, 10.0, size=shape)
    _add_sources(image, rng, 200)
    image[rng.random(shape) < 0.01] = np.nan
    return image.astype(np.float32)


def sky_gradient_f64(shape=(4096, 4096), seed: int = BASE_SEED + 2) -> np.ndarray:
    """31x31 workload: N(1000,10) + x-gradient 0..200, 500 sources, 0.1% hot
    pixels U(1e4, 1e5), 2% NaN and 3 fully-NaN columns, float64."""
    rng = np.random.default_rng(seed)
    rows, cols = shape
    image = rng.normal(1000.0, 10.0, size=shape)
    image += np.linspace(0.0, 200.0, cols)[None, :]
    _add_sources(image, rng, 500)
    hot = rng.random(shape) < 0.001
    image[hot] = rng.uniform(1e4, 1e5, size=int(hot.sum()))
    image[rng.random(shape) < 0.02] = np.nan
    image[:, rng.choice(cols, size=3, replace=False)] = np.nan
    return image


def sky_gradient_crop(shape=(1024, 1024), seed: int = BASE_SEED + 7, field=(4096, 4096)) -> np.ndarray:
    """A crop of the 31x31 workload's law: the x-gradient (0..200 across the
    4096-column field) seen through a window of `shape` at a seeded random
    offset, the field's source density (500 per 4096^2), 0.1% hot pixels
    U(1e4, 1e5), 2% NaN and those of the field's 3 NaN columns that fall in
    the crop, float64."""
    rng = np.random.default_rng(seed)
    rows, cols = shape
    oy = int(rng.integers(0, field[0] - rows + 1))
    ox = int(rng.integers(0, field[1] - cols + 1))
    image = rng.normal(1000.0, 10.0, size=shape)
    image += np.linspace(0.0, 200.0, field[1])[None, ox : ox + cols]
    sources = int(rng.poisson(500.0 * rows * cols / (field[0] * field[1])))
    _add_sources(image, rng, sources)
    hot = rng.random(shape) < 0.001
    image[hot] = rng.uniform(1e4, 1e5, size=int(hot.sum()))
    image[rng.random(shape) < 0.02] = np.nan
    for column in rng.choice(field[1], size=3, replace=False):
        if ox <= column < ox + cols:
            image[:, column - ox] = np.nan
    return image


def sky_annulus(shape=(8192, 8192), seed: int = BASE_SEED + 3) -> np.ndarray:
    """Annulus workload: Poisson(200) sky as float32, 5% of 16x16 blocks NaN,
    0.01% +inf."""
    rng = np.random.default_rng(seed)
    image = rng.poisson(200.0, size=shape).astype(np.float32)
    by, bx = shape[0] // 16, shape[1] // 16
    blocks = rng.random((by, bx)) < 0.05
    image[: by * 16, : bx * 16][np.repeat(np.repeat(blocks, 16, 0), 16, 1)] = np.nan
    image[rng.random(shape) < 0.0001] = np.inf
    return image


def cube(shape=(128, 1024, 1024), seed: int = BASE_SEED + 4) -> np.ndarray:
    """3-D workload: Poisson(200) float32 cube with the annulus NaN/inf laws."""
    rng = np.random.default_rng(seed)
    image = rng.poisson(200.0, size=shape).astype(np.float32)
    image[rng.random(shape) < 0.05] = np.nan
    image[rng.random(shape) < 0.0001] = np.inf
    return image


def sky_quadratic(shape=(8192, 8192), seed: int = BASE_SEED + 5, sources: int = 2000) -> np.ndarray:
    """63x63 workload: N(1000,10) sky + 2-D quadratic background up to 300,
    2000 sources, 2% NaN, float32."""
    rng = np.random.default_rng(seed)
    rows, cols = shape
    image = rng.normal(1000.0, 10.0, size=shape)
    y = np.linspace(-1.0, 1.0, rows)[:, None]
    x = np.linspace(-1.0, 1.0, cols)[None, :]
    image += 150.0 * (x**2 + y**2)
    _add_sources(image, rng, sources)
    image[rng.random(shape) < 0.02] = np.nan
    return image.astype(np.float32)


def sky_quadratic_crop(shape=(2048, 2048), seed: int = BASE_SEED + 6, field=(8192, 8192)) -> np.ndarray:
    """A crop of the 63x63 workload's law: the quadratic background of an
    8192^2 field (up to 300 at its corners) seen through a window of `shape`
    at a seeded random offset, with the field's source density
    (2000 per 8192^2) and 2% NaN, float32. Per-pixel statistics (noise,
    background slope, sources) match the full workload image."""
    rng = np.random.default_rng(seed)
    rows, cols = shape
    oy = int(rng.integers(0, field[0] - rows + 1))
    ox = int(rng.integers(0, field[1] - cols + 1))
    image = rng.normal(1000.0, 10.0, size=shape)
    y = np.linspace(-1.0, 1.0, field[0])[oy : oy + rows, None]
    x = np.linspace(-1.0, 1.0, field[1])[None, ox : ox + cols]
    image += 150.0 * (x**2 + y**2)
    sources = int(rng.poisson(2000.0 * rows * cols / (field[0] * field[1])))
    _add_sources(image, rng, sources)
    image[rng.random(shape) < 0.02] = np.nan
    return image.astype(np.float32)


def rank_for_quantile(quantile: float, count: int) -> int:
    """Rank at a quantile over the full tap count: floor(q * (W - 1) + 0.5)."""
    return int(np.floor(quantile * (count - 1) + 0.5))


def cutouts(count: int = 512, seed: int = BASE_SEED + 8):
    """Batched-cutout workload: `count` float32 images with independent
    log-uniform side lengths in [64, 1024], each with its own odd window size
    in {3, 5, ..., 15} and boundary mode, sky N(100, 5) plus 1% NaN.
    Returns (images, sizes, modes)."""
    rng = np.random.default_rng(seed)
    modes_all = ("reflect", "constant", "nearest", "mirror", "wrap")
    images, sizes, modes = [], [], []
    for _ in range(count):
        rows, cols = (int(v) for v in np.exp(rng.uniform(np.log(64.0), np.log(1024.0), size=2)))
        image = rng.normal(100.0, 5.0, size=(rows, cols))
        image[rng.random((rows, cols)) < 0.01] = np.nan
        images.append(image.astype(np.float32))
        sizes.append(int(rng.choice(np.arange(3, 16, 2))))
        modes.append(modes_all[int(rng.integers(0, 5))])
    return images, sizes, modes
