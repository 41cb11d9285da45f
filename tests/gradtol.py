"""Tolerance rule for point-distance adjoints against the FP64 closed form.

For d_ij = |q_i - p_j| the kernel forms q - p from two float32 inputs, which
is correctly rounded to 2^-24 of the difference itself (exact when the two
are within a factor of two, Sterbenz), so each neighbour's term
g_ij (q - p)/d carries a few float32 roundings of its own size and nothing
that scales with |q|.  The bound is therefore relative to the gradient's own
magnitude:

* per query:  |gq_i - gq64_i| <= 1e-5 * sum_j |g_ij|   (every component);
* per point:  |gp_p - gp64_p| <= max(1e-5, (n_p + 4) * 2^-24) * sum |g|
  over the n_p terms scattered into p (float32 accumulation of n_p terms in
  any order is within (n_p - 1) * 2^-24 of their absolute sum).

A sign flip or a wrong neighbour changes a query's gradient by the size of
that neighbour's term, far above 1e-5 of the sum.
"""
import numpy as np

EPS = 2.0 ** -24


def check_point_grads(oracle_mod, P, Q, ids, g, gq, gp=None, label="", gp_base=None):
    """Asserts the rule above for kernel gradients gq [nq,3] (and gp [np,3],
    accumulated onto gp_base, default zeros) and prints the max and p99 of
    the relative error.  ids and g are [nq] or [nq,k]."""
    ids = np.asarray(ids).reshape(len(Q), -1)
    g = np.asarray(g, dtype=np.float64).reshape(len(Q), -1)
    g64q, g64p = oracle_mod.point_grad_f64(P, Q, ids, g)
    valid = ids >= 0
    scale_q = (np.abs(g) * valid).sum(1)
    err_q = np.abs(np.asarray(gq, dtype=np.float64) - g64q).max(1)
    rel_q = err_q / np.maximum(scale_q, 1e-300)
    rel_q[scale_q == 0] = 0.0
    print(f"[grad] {label} gradQ rel err (of sum|g|): max {rel_q.max():.2e} "
          f"p99 {np.quantile(rel_q, 0.99):.2e}", flush=True)
    assert np.all(err_q <= 1e-5 * scale_q), f"{label}: query gradient off by {rel_q.max():.2e}"
    if gp is None:
        return rel_q
    flat_ids = ids[valid]
    scale_p = np.bincount(flat_ids, np.abs(g[valid]), minlength=len(P))
    n_p = np.bincount(flat_ids, minlength=len(P))
    if gp_base is not None:
        g64p = g64p + gp_base
    err_p = np.abs(np.asarray(gp, dtype=np.float64) - g64p).max(1)
    bound = np.maximum(1e-5, (n_p + 4) * EPS) * scale_p
    rel_p = err_p / np.maximum(scale_p, 1e-300)
    rel_p[scale_p == 0] = 0.0
    print(f"[grad] {label} gradP rel err (of sum|g|): max {rel_p.max():.2e} "
          f"p99 {np.quantile(rel_p, 0.99):.2e} (max terms per point {n_p.max()})", flush=True)
    assert np.all(err_p <= bound), f"{label}: point gradient off by {rel_p.max():.2e}"
    return rel_q, rel_p


def check_chamfer_grads(oracle_mod, S, T, nn_s, nn_t, grad_s, grad_t, label=""):
    """One cloud pair of the batched Chamfer adjoint: loss = mean_i |s_i -
    T[nn_s[i]]| + mean_j |t_j - S[nn_t[j]]|.  Source point i collects its
    own query term (weight 1/|S|) and the data terms of every target query
    whose neighbour it is (weight 1/|T| each); likewise for targets.  The
    bound per point is the per-point rule above over all of its terms."""
    ns, nt = len(S), len(T)
    ws, wt = np.full(ns, 1.0 / ns), np.full(nt, 1.0 / nt)
    gq_s, gp_t = oracle_mod.point_grad_f64(T, S, nn_s, ws)
    gq_t, gp_s = oracle_mod.point_grad_f64(S, T, nn_t, wt)
    worst = 0.0
    for name, g, g64, own, other_ids, other_w, n in [
            ("src", grad_s, gq_s + gp_s, ws * (nn_s >= 0), nn_t, wt, ns),
            ("tgt", grad_t, gq_t + gp_t, wt * (nn_t >= 0), nn_s, ws, nt)]:
        ok = other_ids >= 0
        scale = own + np.bincount(other_ids[ok], other_w[ok], minlength=n)
        terms = (own > 0) + np.bincount(other_ids[ok], minlength=n)
        err = np.abs(np.asarray(g, dtype=np.float64) - g64).max(1)
        bound = np.maximum(1e-5, (terms + 4) * EPS) * scale
        rel = np.where(scale > 0, err / np.maximum(scale, 1e-300), 0.0)
        worst = max(worst, float(rel.max()) if n else 0.0)
        assert np.all(err <= bound), f"{label} grad {name}: rel err {rel.max():.2e}"
    return worst
