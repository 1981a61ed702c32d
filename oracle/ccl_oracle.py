"""CPU oracle of the bubble / string component analysis -- TEST INFRASTRUCTURE ONLY.

Only tests/ (and the benchmark's CPU-baseline leg) use this module, as the checker.  It
restates postprocess/revisualize_snapshots.py:171-213 (_identify_bubbles, mask of :248-250)
and :360-392 (_identify_strings) on top of scipy.ndimage.label, the routine the reference
calls (scipy is installed here), and holds an independent breadth-first labeller that pins
scipy's connectivity and numbering (and provides the periodic variant scipy lacks).
"""
from __future__ import annotations

from collections import deque

import numpy as np
from scipy import ndimage


def label_bfs(mask, periodic=False):
    """Breadth-first 6-connected labelling; components numbered 1..n in raster order of their
    first site (pure Python: small inputs only)."""
    mask = np.asarray(mask, dtype=bool)
    nx, ny, nz = mask.shape
    lab = np.zeros(mask.shape, np.int32)
    n = 0
    for s in range(mask.size):
        i, j, k = np.unravel_index(s, mask.shape)
        if not mask[i, j, k] or lab[i, j, k]:
            continue
        n += 1
        lab[i, j, k] = n
        q = deque([(i, j, k)])
        while q:
            a, b, c = q.popleft()
            for da, db, dc in ((1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1),
                               (0, 0, -1)):
                p, r, t = a + da, b + db, c + dc
                if periodic:
                    p, r, t = p % nx, r % ny, t % nz
                elif not (0 <= p < nx and 0 <= r < ny and 0 <= t < nz):
                    continue
                if mask[p, r, t] and not lab[p, r, t]:
                    lab[p, r, t] = n
                    q.append((p, r, t))
    return lab, n


def label(mask):
    """scipy.ndimage.label with the default structure (what the reference calls)."""
    return ndimage.label(np.asarray(mask, dtype=bool))


def identify_bubbles(phi, esc, dx_phys=None):
    """revisualize_snapshots.py:171-213 on the mask (phi > esc) | (phi < -esc) (:248-250)."""
    mask = (phi > esc) | (phi < -esc)
    labelled, n_clusters = ndimage.label(mask)
    bubbles = []
    for cid in range(1, n_clusters + 1):
        coords = np.argwhere(labelled == cid)
        vol = len(coords)
        centroid = coords.mean(axis=0)
        r_eff = (3.0 * vol / (4.0 * np.pi)) ** (1.0 / 3.0)
        r_rms = np.sqrt(np.mean(np.sum((coords - centroid) ** 2, axis=1)))
        vals = phi[labelled == cid]
        entry = dict(centroid=centroid, volume=vol, r_eff=r_eff, r_rms=r_rms,
                     sign=1 if np.mean(vals) >= 0 else -1, peak=float(np.max(np.abs(vals))))
        if dx_phys is not None:
            entry["r_phys"] = r_eff * dx_phys
        bubbles.append(entry)
    bubbles.sort(key=lambda b: b["volume"], reverse=True)
    return bubbles


def identify_strings(winding, threshold=0.5):
    """revisualize_snapshots.py:360-392."""
    mask = np.abs(winding) > threshold
    labelled, n_loops = ndimage.label(mask)
    if n_loops == 0:
        return labelled, []
    labels = np.arange(1, n_loops + 1)
    counts = np.bincount(labelled.ravel(), minlength=n_loops + 1)[1:]
    centers = ndimage.center_of_mass(mask, labelled, labels)
    means = ndimage.mean(winding, labelled, labels)
    maxes = ndimage.maximum(np.abs(winding), labelled, labels)
    strings = []
    for i, lid in enumerate(labels):
        strings.append(dict(loop_id=int(lid), n_voxels=int(counts[i]),
                            centroid=np.asarray(centers[i], dtype=np.float64),
                            winding_sign=1 if float(means[i]) >= 0 else -1,
                            max_winding=float(maxes[i])))
    strings.sort(key=lambda s: s["n_voxels"], reverse=True)
    return labelled, strings
