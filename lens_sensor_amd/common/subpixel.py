"""Sub-pixel permittivity smoothing: volume fractions of beads and 1-D layer averages.

Same definitions as common/subpixel.py of the reference (supersampling is the
definition of the fraction, so the sample positions and the inside test are kept):
  sample offsets          common/subpixel.py:21-22
  circle_fraction (S=8)   common/subpixel.py:25-39
  sphere_fraction (S=6)   common/subpixel.py:42-62
  blend                   common/subpixel.py:65-69
  cell_average_column     common/subpixel.py:80-94 (S=16)
Only cells whose centre lies within half a cell diagonal of the surface are sampled;
all others are exactly 0 or 1.
"""
import numpy as np

LAYER_SUBSAMPLES = 16


def _sample_offsets(S, h):
    return ((np.arange(S) + 0.5) / S - 0.5) * h


def circle_fraction(xs, ys, xc, yc, r, h, S=8):
    """Area fraction of the disc (xc, yc, r) in the cells centred on xs x ys."""
    X = xs[:, None] - xc
    Y = ys[None, :] - yc
    dist = np.sqrt(X * X + Y * Y)
    half_diag = h * np.sqrt(2.0) / 2.0
    frac = (dist <= r - half_diag).astype(np.float64)
    ii, jj = np.nonzero((dist > r - half_diag) & (dist < r + half_diag))
    if len(ii):
        o = _sample_offsets(S, h)
        sx = X[ii, 0][:, None, None] + o[None, :, None]
        sy = Y[0, jj][:, None, None] + o[None, None, :]
        frac[ii, jj] = np.mean((sx * sx + sy * sy) <= r * r, axis=(1, 2))
    return frac


def sphere_fraction(xs, ys, zs, xc, yc, zc, r, h, S=6):
    """Volume fraction of the sphere (xc, yc, zc, r) in the cells centred on xs x ys x zs."""
    X = xs[:, None, None] - xc
    Y = ys[None, :, None] - yc
    Z = zs[None, None, :] - zc
    dist = np.sqrt(X * X + Y * Y + Z * Z)
    half_diag = h * np.sqrt(3.0) / 2.0
    frac = (dist <= r - half_diag).astype(np.float64)
    ii, jj, kk = np.nonzero((dist > r - half_diag) & (dist < r + half_diag))
    if len(ii):
        o = _sample_offsets(S, h)
        inside = np.zeros(len(ii))
        sy = Y[0, jj, 0][:, None, None] + o[None, :, None]
        sz = Z[0, 0, kk][:, None, None] + o[None, None, :]
        for ox in o:                      # one x-slice of samples at a time
            sx = (X[ii, 0, 0] + ox)[:, None, None]
            inside += np.sum((sx * sx + sy * sy + sz * sz) <= r * r, axis=(1, 2))
        frac[ii, jj, kk] = inside / S ** 3
    return frac


def blend(eps, frac, n_incl):
    """eps <- (1 - f) eps + f n^2 where f > 0 (in place on the given view)."""
    m = frac > 0
    eps[m] = (1.0 - frac[m]) * eps[m] + frac[m] * (n_incl ** 2)
    return eps


def cell_average_column(ys, dy, eps_of_y, S=LAYER_SUBSAMPLES, harmonic=False, shift=0.0):
    """Average of eps(y) over the cell [y_j + shift - dy/2, y_j + shift + dy/2].
    Arithmetic mean for field components parallel to the layers, harmonic mean
    (``harmonic=True``) for the component normal to them."""
    off = ((np.arange(S) + 0.5) / S - 0.5) * dy + shift
    Ys = np.asarray(ys, dtype=np.float64)[:, None] + off[None, :]
    e = np.asarray(eps_of_y(Ys.ravel()))
    is_complex = np.iscomplexobj(e)
    e = e.astype(np.complex128 if is_complex else np.float64).reshape(Ys.shape)
    out_dtype = np.complex64 if is_complex else np.float32
    if harmonic:
        return (1.0 / np.mean(1.0 / e, axis=1)).astype(out_dtype)
    return np.mean(e, axis=1).astype(out_dtype)
