"""Shared checks for the parity tests."""
import numpy as np


def classify_hits(cpu, gpu, rel_t=1e-5, edge_eps=1e-4):
    """Compares closest-hit records.  A ray agrees when both miss, or the triangle ids
    match; a mismatch is allowed when the hit distances tie (|dt| <= rel_t * t) or
    the CPU hit lies on a triangle edge/vertex (a barycentric within edge_eps of 0).
    Returns a dict of counts and the t error of agreeing rays."""
    cp, gp = cpu["prim"], gpu["prim"]
    ct, gt = cpu["t"].astype(np.float64), gpu["t"].astype(np.float64)
    both_miss = (cp < 0) & (gp < 0)
    same = (cp == gp) & (cp >= 0)
    hit_both = (cp >= 0) & (gp >= 0)
    # distances are compared only where both rays hit (a miss is t = inf)
    dt = np.abs(np.where(hit_both, ct, 0.0) - np.where(hit_both, gt, 0.0))
    tie = hit_both & ~same & (dt <= rel_t * np.abs(np.where(hit_both, ct, 0.0)))
    b0 = 1.0 - cpu["b1"] - cpu["b2"]
    edge = (np.minimum(np.minimum(b0, cpu["b1"]), cpu["b2"]) <= edge_eps) & (cp >= 0)
    gb0 = 1.0 - gpu["b1"] - gpu["b2"]
    gedge = (np.minimum(np.minimum(gb0, gpu["b1"]), gpu["b2"]) <= edge_eps) & (gp >= 0)
    agree = both_miss | same
    explained = ~agree & (tie | edge | gedge)
    fail = ~agree & ~explained
    terr = np.abs(ct[same] - gt[same]) / np.maximum(np.abs(ct[same]), 1e-30)
    return {"n": len(cp), "agree": int(agree.sum()), "explained": int(explained.sum()),
            "fail": int(fail.sum()), "max_rel_t": float(terr.max()) if terr.size else 0.0,
            "fail_idx": np.nonzero(fail)[0]}
