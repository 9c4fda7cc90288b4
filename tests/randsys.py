"""Seeded random small systems shared by the randomised parity tests."""
import numpy as np

from tcp_allpair_amd.inputs import make_plasma
from tcp_allpair_amd.params import RECIP_NONE


def random_system(seed, nmax=48):
    """A random system as the reference's driver could read it, up to nmax
    particles of each species; seeds 0 mod 7 have no electrons, 3 mod 7 no
    ions."""
    rng = np.random.default_rng(1000 + seed)
    ne = 0 if seed % 7 == 0 else int(rng.integers(1, nmax))
    ni = 0 if seed % 7 == 3 else int(rng.integers(1, nmax))
    ne, ni = max(ne, 2) if ni == 0 else ne, max(ni, 2) if ne == 0 else ni
    p = make_plasma(1, 1, recip=RECIP_NONE)[0]
    L = ((ne + ni) / rng.uniform(0.3, 3.0)) ** (1.0 / 3.0)
    box = tuple(L * rng.uniform(0.7, 1.4, 3))
    rc = rng.uniform(0.25, 0.5) * min(box)
    nbins = int(rng.choice([120, 501, 900]))  # 900: pairs past bin Ndist+1
    p = p.replace(ne=ne, ni=ni, box=box, rc=rc, rs=rng.uniform(0.2, 1.0),
                  dx=tuple(rc / nbins * rng.uniform(0.8, 1.2, 3)), tau=rng.uniform(0.02, 0.5),
                  Te=p.Te * rng.uniform(0.5, 2.0), Ti=p.Ti * rng.uniform(0.5, 2.0))
    x = (rng.random((ne + ni, 3)) - 0.5) * np.array(box)
    sig = np.sqrt(np.where(np.arange(ne + ni) < ne, p.Te / p.xm[0], p.Ti / p.xm[1]))[:, None]
    v = rng.standard_normal((ne + ni, 3)) * sig
    # particles exactly on box faces, and one pair a cut-off apart
    for i in range(min(3, ne + ni)):
        x[i, i % 3] = (0.5 if i % 2 else -0.5) * box[i % 3]
    if ne + ni >= 5:
        x[4] = x[3] + np.array([rc, 0.0, 0.0])
        x[4, 0] -= box[0] * np.round(x[4, 0] / box[0])
    return p, x, v
