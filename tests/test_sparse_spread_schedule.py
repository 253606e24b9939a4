"""CPU model of the sparse plane spread's schedule (k_pme_spec2d, many-site
solutes, in ermod-openacc_amd/csrc/er_kernels.h): one wave adds four sites'
ORD x ORD stencil blocks per step, lane (k, p) taking point p of site
j + k Q of group j (Q = a quarter of the chunk).  A group whose blocks meet
goes in two steps (k = 0, 2 then 1, 3) when those pairs do not meet, else in
k order.  Checked here: no two lanes of one step add to the same plane point
(the adds of a step are one LDS read-modify-write per lane), every (site,
point) is added exactly once, and the meet test agrees with brute force."""
import numpy as np


def blocks_meet(a, b, K2, K3, order):
    """The kernel's test: circular distance of the block corners below the
    order on both axes."""
    dx, dy = (a[0] - b[0]) % K2, (a[1] - b[1]) % K3
    return min(dx, K2 - dx) < order and min(dy, K3 - dy) < order


def block(g, K2, K3, order):
    return {((g[0] - c2) % K2, (g[1] - c3) % K3) for c2 in range(order) for c3 in range(order)}


def schedule(g, K2, K3, order):
    """[[(site, point), ...] per step] as the kernel issues them."""
    ns = len(g)
    Q = (ns + 3) // 4
    steps = []
    for j in range(Q):
        sites = [j + u * Q if j + u * Q < ns else None for u in range(4)]
        pairs = [(u, v) for u in range(4) for v in range(u + 1, 4) if sites[u] is not None and sites[v] is not None]
        meet = any(blocks_meet(g[sites[u]], g[sites[v]], K2, K3, order) for u, v in pairs)
        far = any(blocks_meet(g[sites[u]], g[sites[v]], K2, K3, order) for u, v in pairs if v == u + 2)
        lanes = lambda ks: [(sites[k], pt) for k in ks if sites[k] is not None for pt in block(g[sites[k]], K2, K3, order)]
        if not meet:
            steps.append(lanes(range(4)))
        elif not far:
            steps += [lanes((0, 2)), lanes((1, 3))]
        else:
            steps += [lanes((k,)) for k in range(4)]
    return steps


def test_meet_test_is_exact():
    rng = np.random.default_rng(11)
    for _ in range(4000):
        K2, K3 = (int(x) for x in rng.choice([16, 20, 32, 72, 108, 128], 2))
        order = int(rng.choice([3, 4]))
        a, b = rng.integers(0, K2), rng.integers(0, K3)
        c, d = rng.integers(0, K2), rng.integers(0, K3)
        share = bool(block((a, b), K2, K3, order) & block((c, d), K2, K3, order))
        assert share == blocks_meet((a, b), (c, d), K2, K3, order)


def test_steps_are_race_free_and_complete():
    rng = np.random.default_rng(5)
    n_two = n_four = 0
    for trial in range(400):
        K2, K3 = (int(x) for x in rng.choice([16, 32, 72, 108, 128], 2))
        order = int(rng.choice([3, 4]))
        ns = int(rng.integers(1, 160))
        if trial % 2:   # a chain: bonded sites a grid point or two apart
            g = np.cumsum(rng.integers(-2, 3, (ns, 2)), 0) + rng.integers(0, 128, 2)
        else:
            g = rng.integers(0, 128, (ns, 2))
        g = [(int(x) % K2, int(y) % K3) for x, y in g]
        steps = schedule(g, K2, K3, order)
        n_two += sum(1 for s in steps if s and len({site for site, _ in s}) == 2)
        n_four += sum(1 for s in steps if s and len({site for site, _ in s}) == 1)
        done = []
        for s in steps:
            pts = [pt for _, pt in s]
            assert len(pts) == len(set(pts)), (trial, s[:4])   # one lane per point in a step
            done += s
        want = sorted((site, pt) for site in range(ns) for pt in block(g[site], K2, K3, order))
        assert sorted(done) == want
    assert n_two > 0 and n_four > 0   # both fallback schedules were exercised
