"""Helpers comparing GPU hit streams with the oracle's intersections."""
import numpy as np

MISS = 0xFFFFFFFF


def compare_hits(gpu_hits, ids, inter, occlusion=False):
    """Returns a dict of agreement statistics between GPU hits (HitBuffers)
    and oracle intersections (INTER_DTYPE) for the same rays."""
    g_inst = gpu_hits.inst_u32()
    g_t = gpu_hits.t.cpu().numpy()
    g_u = gpu_hits.u.cpu().numpy()
    g_v = gpu_hits.v.cpu().numpy()
    g_hit = g_inst != MISS
    o_hit = inter["hit"] != 0
    n = len(o_hit)
    hit_agree = g_hit == o_hit
    both = g_hit & o_hit
    gid = np.zeros(n, np.uint64)
    gid[g_hit] = ids[g_inst[g_hit]]
    id_agree = np.ones(n, bool)
    id_agree[both] = gid[both] == inter["id"][both]
    out = {"n": n, "hits": int(o_hit.sum()), "hit_agree": float(hit_agree.mean()),
           "id_agree": float(id_agree.mean())}
    if not occlusion:
        ot = inter["t"]
        rel = np.zeros(n)
        rel[both] = np.abs(g_t[both] - ot[both]) / np.maximum(np.abs(ot[both]), 1e-30)
        exact = np.ones(n, bool)
        exact[both] = g_t[both].view(np.uint32) == ot[both].view(np.uint32)
        out["t_exact"] = float(exact.mean())
        out["t_rel_max"] = float(rel.max()) if n else 0.0
        ok = hit_agree & id_agree & (rel <= 1e-5)
        out["match"] = float(ok.mean())
        same = both & exact & id_agree
        out["uv_exact_on_same_t"] = float(((g_u[same] == inter["u"][same]) & (g_v[same] == inter["v"][same])).mean()) \
            if same.any() else 1.0
    else:
        out["match"] = float(hit_agree.mean())
    return out
