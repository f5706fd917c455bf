"""Draws from the whole BC1 encoder settings space, shared by the CPU and GPU
parity tests.

The reference exposes every field EncodeBlock reads as a writable property
(quicktex/s3tc/bc1/_bindings.cpp:134-164): error mode, endpoint mode, the three
two-pass flags, exhaustive, search rounds, orderings and power iterations.  The
presets of SetLevel (BC1Encoder.cpp:74-249) cover 20 points of that space; these
draws cover the rest, in strata chosen to reach every branch of EncodeBlock
(BC1Encoder.cpp:261-342) and both builds of the queued lane kernel:

* ``free``: every field drawn independently, raw values past the setters'
  clamps included (orderings 0 and above the table, power iterations outside
  4..10: BC1Encoder.cpp:251-258);
* ``none``: ErrorMode::None (BC1Encoder.cpp:276-278, 695-720), whose block error
  stays 0 so that no refinement runs;
* ``hq``: the settings that select the level 13-18 kernel build (Full error,
  PCA endpoints, two least-squares and two endpoint passes, no exhaustive fit)
  with every other field drawn;
* ``edges``: search rounds at the exit boundaries 0, 1, 33, 34, 35, 100, 256
  and orderings at both ends of their range;
* ``light``: the settings that select the level 0-12 kernel builds (no endpoint
  search, one endpoint and one cluster-fit pass, no exhaustive fit) with every
  other field drawn.

Each draw is a raw ``BC1Params``: callers clamp it the way their encoder does
(the product's qtx_bc1_params_clamp, the reference's setters).
"""
import numpy as np

from quicktex_amd.capi import BC1Params

ROUNDS = (0, 1, 33, 34, 35, 100, 256)
MODES = (4, 3, 0x13)


def _draw(rng, stratum, exhaustive_ok):
    p = BC1Params()
    p.interpolator = int(rng.integers(0, 4))
    p.color_mode = int(rng.choice(MODES))
    p.error_mode = int(rng.integers(0, 4))
    p.endpoint_mode = int(rng.integers(0, 4))
    p.two_ls_passes = int(rng.integers(0, 2))
    p.two_ep_passes = int(rng.integers(0, 2))
    p.two_cf_passes = int(rng.integers(0, 2))
    p.exhaustive = int(exhaustive_ok and rng.random() < 0.15)
    p.power_iterations = int(rng.choice([0, 3, 4, 5, 6, 7, 8, 9, 10, 11, 1000]))
    p.search_rounds = int(rng.choice(ROUNDS)) if rng.random() < 0.7 else int(rng.integers(0, 300))
    p.orderings4 = int(rng.choice([0, 1, 2, 3, 16, 31, 64, 100, 127, 128, 129, 5000]))
    p.orderings3 = int(rng.choice([0, 1, 2, 7, 15, 16, 31, 32, 33, 900]))
    if stratum == 'none':
        p.error_mode = 0
    elif stratum == 'hq':
        p.error_mode, p.endpoint_mode, p.two_ls_passes, p.two_ep_passes, p.exhaustive = 3, 3, 1, 1, 0
    elif stratum == 'light':
        p.search_rounds, p.two_ep_passes, p.two_cf_passes, p.exhaustive = 0, 0, 0, 0
    elif stratum == 'edges':
        p.search_rounds = int(rng.choice(ROUNDS))
        p.orderings4 = int(rng.choice([1, 128]))
        p.orderings3 = int(rng.choice([1, 32]))
    return p


def draws(seed, n_per_stratum, exhaustive_ok=True):
    """[(name, raw BC1Params)] -- n_per_stratum draws of every stratum, plus the
    16 combinations of error mode x endpoint mode with the other fields drawn."""
    rng = np.random.default_rng(seed)
    out = []
    for stratum in ('free', 'none', 'hq', 'edges', 'light'):
        for k in range(n_per_stratum):
            out.append(('%s%d' % (stratum, k), _draw(rng, stratum, exhaustive_ok)))
    for em in range(4):
        for ep in range(4):
            p = _draw(rng, 'free', exhaustive_ok)
            p.error_mode, p.endpoint_mode = em, ep
            out.append(('E%dP%d' % (em, ep), p))
    return out


def describe(p):
    return ('interp=%d mode=%#x err=%d ep=%d ls2=%d ep2=%d cf2=%d ex=%d pi=%d rounds=%d o4=%d o3=%d' %
            p.as_tuple())


def copy(p):
    q = BC1Params()
    for name, _ in BC1Params._fields_:
        setattr(q, name, getattr(p, name))
    return q
