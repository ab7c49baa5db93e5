"""Shared helpers for the parity tests: the geometry epoch restated through the oracle."""
from __future__ import annotations

import numpy as np


def root_cube(info):
    w = np.float32(2.0002)*np.float32(info["y_max"])
    o = np.float32(-0.5)*w
    return (o, o, o), w


def oracle_geometry(o, pts, act, q, y, z):
    """Morton keys, sort, hierarchy, tree top-K and measures from given sites (particle order)."""
    info = o.info
    org, w = root_cube(info)
    keys = o.morton_keys(pts, org, w, info["morton_max_level"], info["dim"])
    perm = o.sort_perm(keys)
    H = o.hierarchy(keys[perm], org, w, info["dim"], info["morton_leaf_target"], info["morton_max_level"])
    n = len(pts)
    codes, d2 = o.topk_tree(pts, act, perm, H, info["dim"], info["morton_max_level"], np.arange(n, dtype=np.int32),
                            info["n_k"], q)
    codes[act == 0] = -1
    valid = codes >= 0
    maxd2 = np.where(valid, d2, np.float32(0)).max(axis=1)
    meas = np.zeros(n)
    for i in np.nonzero(act)[0]:
        meas[i] = o.ball_measure(float(y[i]), float(z[i]), float(np.sqrt(np.float64(maxd2[i]))))
    return {"keys": keys, "perm": perm, "nodes": H, "codes": codes, "d2": d2, "measure": meas}
