"""CPU oracle for the static elasticity solve: a NumPy restatement of the reference.

THIS MODULE IS TEST INFRASTRUCTURE (same rule as ``oracle/solver.py``): only
``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline`` leg
may import it, as the checker.

``elastic_solve_static`` restates ``src/shearwave/solver.py:569-625``: CG on
``(-mu lap) u = bz`` with float32 vectors, float32 scalars and NumPy's float32
``np.sum`` for the dot products, exactly as the reference runs it.

``numpy_sum_f32`` restates the order in which ``np.sum`` adds a C-contiguous
float32 array (NumPy 2.2.6, the version ``uv.lock:905-906`` pins for the
reference): the array is taken in buffers of 8192 elements, the running total
(starting at 0) adds each buffer's pairwise sum in order, and the pairwise sum
of ``n`` elements is a plain loop below 8 elements, eight interleaved
accumulators up to 128 elements (combined as ((r0+r1)+(r2+r3))+((r4+r5)+(r6+r7))
and followed by the leftover elements in order), and above that the sums of
the first ``n2 = n//2 - (n//2) % 8`` elements and of the rest. ``pairwise_plan``
lists the same tree as leaves plus internal nodes level by level, the form in
which the GPU evaluates it.

Pinning (tests/test_oracle_static_solve.py): ``numpy_sum_f32`` and the plan
against ``np.sum`` itself bit for bit over many sizes and shapes; the operator
against an explicit matrix; the solve against the reference's own
manufactured-solution test (test_projection_static.py:86-105), against a
sparse direct solve of the same discrete system, and on the sign convention
the reference's docstring records (a non-zero solution).
"""

from __future__ import annotations

import numpy as np

from .solver import interior_mask, laplacian_center, prepare_material_field

F32_EPS = np.finfo(np.float32).eps
NPY_BUFSIZE = 8192  # np.getbufsize() default
PW_BLOCKSIZE = 128  # NumPy's pairwise leaf size


# ---------------------------------------------------------------------------
# NumPy's float32 summation order


def _split(n):
    n2 = n // 2
    return n2 - n2 % 8


def _leaf_sum(a):
    """One pairwise leaf (n <= 128) in float32, in NumPy's order."""
    n = len(a)
    if n < 8:
        res = np.float32(0.0)
        for v in a:
            res = np.float32(res + v)
        return res
    full = n - n % 8
    r = a[:8].copy()
    for o in range(8, full, 8):
        r = r + a[o : o + 8]  # eight float32 accumulators
    res = np.float32((r[0] + r[1]) + (r[2] + r[3])) + np.float32((r[4] + r[5]) + (r[6] + r[7]))
    res = np.float32(res)
    for v in a[full:]:
        res = np.float32(res + v)
    return res


def pairwise_sum_f32(a):
    """NumPy's pairwise sum of a 1-D float32 array (no buffering)."""
    a = np.asarray(a, dtype=np.float32)
    n = len(a)
    if n <= PW_BLOCKSIZE:
        return _leaf_sum(a)
    if n & (n - 1) == 0:
        # a power of two (> 128): a perfect tree over 128-element leaves, evaluated vectorised
        leaves = a.reshape(-1, 16, 8)
        acc = leaves[:, 0, :].copy()
        for j in range(1, 16):
            acc = acc + leaves[:, j, :]
        v = ((acc[:, 0] + acc[:, 1]) + (acc[:, 2] + acc[:, 3])) + ((acc[:, 4] + acc[:, 5]) + (acc[:, 6] + acc[:, 7]))
        while len(v) > 1:
            v = v[0::2] + v[1::2]
        return np.float32(v[0])
    n2 = _split(n)
    return np.float32(pairwise_sum_f32(a[:n2]) + pairwise_sum_f32(a[n2:]))


def numpy_sum_f32(a):
    """``np.sum`` of a C-contiguous float32 array, restated (buffers of 8192, pairwise within)."""
    flat = np.ascontiguousarray(a, dtype=np.float32).ravel()
    total = np.float32(0.0)
    for s in range(0, len(flat), NPY_BUFSIZE):
        total = np.float32(total + pairwise_sum_f32(flat[s : s + NPY_BUFSIZE]))
    return total


def pairwise_plan(n):
    """The pairwise tree of n <= 8192 elements as (leaves, levels).

    leaves: [(start, length)] in order. levels: lists of (left, right) node ids,
    where ids 0..L-1 are the leaves and internal nodes are numbered L, L+1, ...
    in the order the levels list them (a level's children are all in earlier
    levels). The root is the last internal node, or leaf 0 when L == 1.
    """
    leaves = []
    nodes = []  # (left, right, height), appended in post-order

    def build(start, m):
        if m <= PW_BLOCKSIZE:
            leaves.append((start, m))
            return ("leaf", len(leaves) - 1, 0)
        n2 = _split(m)
        left = build(start, n2)
        right = build(start + n2, m - n2)
        nodes.append((left, right, 1 + max(left[2], right[2])))
        return ("node", len(nodes) - 1, nodes[-1][2])

    build(0, n)
    n_leaf = len(leaves)
    order = sorted(range(len(nodes)), key=lambda k: (nodes[k][2], k))  # by height, stable
    new_id = {k: n_leaf + i for i, k in enumerate(order)}

    def ref(x):
        return x[1] if x[0] == "leaf" else new_id[x[1]]

    levels = []
    for h in sorted({nd[2] for nd in nodes}):
        levels.append([(ref(nodes[k][0]), ref(nodes[k][1])) for k in order if nodes[k][2] == h])
    return leaves, levels


def eval_plan(a, plan):
    """Evaluate a pairwise plan over a (n <= 8192) float32 array."""
    leaves, levels = plan
    val = [_leaf_sum(np.asarray(a[s : s + m], dtype=np.float32)) for s, m in leaves]
    for lev in levels:
        val += [np.float32(val[lft] + val[rgt]) for lft, rgt in lev]
    return val[-1]


# ---------------------------------------------------------------------------
# elastic_solve_static


def apply_static_operator(x, mu_field, dX, dY, dZ):
    """``A x = -mu lap(x)`` (solver.py:595-596)."""
    return -mu_field * laplacian_center(x, dX, dY, dZ)


def elastic_solve_static(bz, dX, dY, dZ, mu, tol=1e-6, max_iters=1500, *, history=None, info=None):
    """CG on ``(-mu lap) u = bz``, zero Dirichlet boundary (solver.py:569-625).

    ``history`` (a list) receives (rs_new, rel) as float32 for every completed
    iteration; ``info`` (a dict) receives iters, relres and the stop reason.
    """
    nx, ny, nz = bz.shape
    mu_field = prepare_material_field(mu, nx, ny, nz)
    rhs = bz.astype(np.float32)
    u = np.zeros_like(rhs, dtype=np.float32)
    mask = interior_mask(rhs.shape)

    def apply_a(x):
        return apply_static_operator(x, mu_field, dX, dY, dZ)

    r = rhs - apply_a(u)
    r *= mask
    p = r.copy()
    rs_old = np.sum(r * r)
    rs0 = max(rs_old, F32_EPS)
    reason, iters, rel = "max_iters", 0, None
    for it in range(max_iters):
        a_p = apply_a(p)
        denom = np.sum(p * a_p)
        if denom <= 0.0:
            reason = "denom"
            break
        alpha = rs_old / denom
        u = (u + alpha * p) * mask
        r = (r - alpha * a_p) * mask
        rs_new = np.sum(r * r)
        rel = np.sqrt(rs_new / rs0)
        iters = it + 1
        if history is not None:
            history.append((rs_new, rel))
        if rel < tol:
            reason = "converged"
            break
        beta = rs_new / max(rs_old, F32_EPS)
        p = (r + beta * p) * mask
        rs_old = rs_new
    if info is not None:
        info.update(iters=iters, relres=None if rel is None else float(rel), reason=reason)
    return u
