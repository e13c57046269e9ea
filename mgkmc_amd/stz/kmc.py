"""compute_rates / select_event / decode_index (mgkmc/stz/kmc.py:4-103) on the GPU.

Selection stream: the reference draws r = np.random.uniform(0, total) inside
numba with numba's own, never-seeded MT19937.  Here the draw comes from
``SELECTION_STREAM`` (numpy RandomState, MT19937 init_genrand + genrand_res53),
seeded with ``seed_selection(seed)``; r = total * u, u = random_sample().
"""
import numpy as np

from ..kmc_device import rates_host, select_host

SELECTION_STREAM = np.random.RandomState()


def seed_selection(seed):
    """Seed the event-selection stream (what a jitted np.random.seed does)."""
    SELECTION_STREAM.seed(seed)


def compute_rates(Q_field, volume, temperature, nu0=1e13, cascade_mode=True):
    """(rates, indices, total_rate) of the events in (x, y, z, m) order."""
    return rates_host(np.asarray(Q_field), volume, temperature, nu0, cascade_mode)


def select_event(rates, total_rate):
    """Index into ``rates`` of the event drawn with r = U(0, total_rate): the
    first i whose sequential running sum reaches r; -1 if total_rate <= 0."""
    if total_rate <= 0:
        return -1
    r = total_rate * SELECTION_STREAM.random_sample()
    return select_host(rates, r)


def decode_index(flat_idx, ny, nz, M):
    """Flat (x, y, z, m) index -> coordinates (integer arithmetic)."""
    m = flat_idx % M
    t = flat_idx // M
    z = t % nz
    t = t // nz
    y = t % ny
    x = t // ny
    return x, y, z, m
