"""Loader of the fixtures written by tests/golden/make_golden.py."""
import os

import numpy as np

from tcp_allpair_amd.params import TcpParams

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    d = dict(np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False))
    p = TcpParams(ne=int(d["p_ne"]), ni=int(d["p_ni"]), box=tuple(d["p_box"]), q=tuple(d["p_q"]),
                  xm=tuple(d["p_xm"]), rc=float(d["p_rc"]), rs=float(d["p_rs"]), dx=tuple(d["p_dx"]),
                  Te=float(d["p_Te"]), Ti=float(d["p_Ti"]), K=tuple(int(k) for k in d["p_K"]),
                  tau=float(d["p_tau"]), fric=float(d["p_fric"]))
    return p, d
