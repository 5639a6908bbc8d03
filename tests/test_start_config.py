"""The tiling of a small configuration onto a larger lattice
(susy_amd/start_config.py), checked on the CPU with the oracle: every local
observable of the tiled links equals the small lattice's (plaquettes, the
gauge and bmass actions per site), the site order is the one the gather
tables use, and the committed start file is a readable MILC file of the
stated size."""
import os

import numpy as np
import pytest

from oracle_lib import Oracle, random_links
from susy_amd import start_config as sc


def test_site_index_is_the_gather_order():
    dims = (2, 4, 2, 6)
    O = Oracle(2, dims=dims)
    xyzt = O.coords()
    idx = sc.site_index(dims)
    lex = xyzt[:, 0] + dims[0] * (xyzt[:, 1] + dims[1] * (xyzt[:, 2] + dims[2] * xyzt[:, 3]))
    assert np.array_equal(idx[lex], np.arange(O.V))


def test_tiled_links_keep_local_observables():
    small, factors = (2, 2, 4, 2), (2, 3, 1, 2)
    big = tuple(s * f for s, f in zip(small, factors))
    kw = dict(lam=2.0, bmass=0.4, G=0.05)
    O = Oracle(3, dims=small, **kw)
    U = random_links(O.link_shape(), 9, 0.3)
    O.set_links(U)
    p_small, a_small = O.plaquette(), O.actions()[:3] / O.V
    O = Oracle(3, dims=big, **kw)
    O.set_links(sc.tile(U, small, factors))
    p_big, a_big = O.plaquette(), O.actions()[:3] / O.V
    assert np.allclose(p_big, p_small, rtol=1e-13)
    assert np.allclose(a_big, a_small, rtol=1e-12)


@pytest.mark.skipif(not os.path.exists(sc.START_FILE), reason="start file not generated")
def test_committed_start_file():
    path = sc.START_FILE
    V = int(np.prod(sc.SMALL_DIMS))
    # 96-byte header + single-precision complex links, 5 NCOL^2 per site
    assert os.path.getsize(path) == 96 + V * 5 * 16 * 8
    O = Oracle(4, dims=sc.SMALL_DIMS, **{k: v for k, v in sc.COUPLINGS.items()
                                         if k in ("PBC", "Nroot", "bmass", "G")},
               lam=sc.COUPLINGS["lam"])
    assert O.load_config(path)[0] == 0
