"""Oracle restatement of lightprop2d/profiles.py (numpy only).  TEST INFRASTRUCTURE.

Each generator follows the reference function at the cited line, including its
quirks: ``gaussian_beam`` shifts ``x``/``y`` in place (profiles.py:69-70) and
``round_hole(R)`` keeps radius R/sqrt(2) (profiles.py:79-80).  Random generators
draw from ``rng`` (default: numpy's legacy global RNG, as the reference does), so
``np.random.seed(s)`` in the reference and ``rng=np.random.RandomState(s)`` here
draw the same stream.

Pinned by the reference's own asserts (tests/test_profiles.py:31-38) and by
closed-form checks in tests/test_oracle.py.
"""
import numpy as np

__all__ = ('plane_wave', 'random_wave', 'random_wave_bin', 'gaussian_beam',
           'round_hole', 'random_round_hole', 'random_round_hole_bin',
           'random_round_hole_phase', 'rectangle_hole', 'square_hole',
           'square_slits', 'binning')


def _rng(rng):
    return np.random if rng is None else rng


def binning(field, binning_order):
    """profiles.py:29-32"""
    return np.kron(field, np.ones((binning_order, binning_order)))


def plane_wave(x, y):
    """profiles.py:35-38"""
    return np.ones((len(y), len(x)))


def random_wave_bin(x, y, binning_order=1, rng=None):
    """profiles.py:41-50"""
    return binning(_rng(rng).randint(
        0, 2, size=(int(len(y) // binning_order), int(len(x) // binning_order))),
        binning_order)


def random_wave(x, y, binning_order=1, rng=None):
    """profiles.py:53-61"""
    return binning(_rng(rng).uniform(
        size=(int(len(y) // binning_order), int(len(x) // binning_order))),
        binning_order)


def gaussian_beam(x, y, A0, rho0, x0=0, y0=0):
    """profiles.py:64-71 (mutates x and y in place, as the reference does)."""
    x -= x0
    y -= y0
    return A0 * np.exp(- (x ** 2 + y ** 2) / rho0 ** 2)


def round_hole(x, y, R, x0=0, y0=0):
    """profiles.py:74-81: exp(-r^2/R^2) >= exp(-1/2), i.e. radius R/sqrt(2)."""
    d = gaussian_beam(x - x0, y - y0, 1, R)
    return d >= 1 / np.exp(0.5)


def random_round_hole_bin(x, y, R, x0=0, y0=0, binning_order=1, rng=None):
    """profiles.py:84-91"""
    field = random_wave_bin(x, y, binning_order, rng=rng)
    field[round_hole(x, y, R, x0, y0) == 0] = 0
    return field


def random_round_hole(x, y, R, x0=0, y0=0, binning_order=1, rng=None):
    """profiles.py:94-101"""
    field = random_wave(x, y, binning_order, rng=rng)
    field[round_hole(x, y, R, x0, y0) == 0] = 0
    return field


def random_round_hole_phase(x, y, R, x0=0, y0=0, binning_order=1, rng=None):
    """profiles.py:104-113: round_hole * exp(i*2*pi*u), u masked to 0 outside."""
    angle = random_round_hole(x, y, R, x0, y0, binning_order, rng=rng) * 2 * np.pi
    phase = np.cos(angle) + 1j * np.sin(angle)
    return round_hole(x, y, R, x0, y0) * phase


def rectangle_hole(x, y, dx, dy, x0=0, y0=0):
    """profiles.py:116-121"""
    return (np.abs(x - x0) < (dx / 2)) & (np.abs(y - y0) < (dy / 2))


def square_hole(x, y, d, x0=0, y0=0):
    """profiles.py:124-128"""
    return rectangle_hole(x, y, d, d, x0, y0)


def square_slits(x, y, d, slits_distance, x0=0, y0=0):
    """profiles.py:131-136"""
    dl = slits_distance / 2
    return square_hole(x, y, d, x0 - dl, y0) | square_hole(x, y, d, x0 + dl, y0)
