"""The schedule of the GPU unwrapper's merge kernel (k_uw_merge_lds,
odtbrain_amd/csrc/odt_prepare.hip) restated step by step in Python and
checked against the oracle's merge sequence (oracle.prepare.merge_groups,
the published rules applied edge by edge) on random edge orders and steps.

What the restatement keeps of the kernel: links instead of member lists
(union by size, a root's own increment 0), 16-bit size words with the
not-a-root and majority codes, batches of edges whose ends are walked to
their roots with path splitting from first links loaded a batch earlier,
walks that see only a random part of the previous batch's merges (size
words and links visible independently, as on the second wavefront), merges
in rounds of edges holding all their claims in a hashed table (roots of the
majority group unclaimed), and the in-order fallback.  The GPU tests compare
the kernel itself with the global-memory merge bit for bit; this test pins
the schedule's equivalence to the sequential rules on the CPU."""
import numpy as np
import pytest

from oracle import prepare as P

NOT_ROOT = 0xFFFF
SAT = 0xFFFE


def grid_edges(H, W):
    """Edge k as (p1, p2): horizontal edges row by row, then vertical ones
    (the kernel's edge numbering)."""
    out = [(y * W + x, y * W + x + 1) for y in range(H) for x in range(W - 1)]
    out += [(p, p + W) for p in range((H - 1) * W)]
    return out


def kernel_schedule(H, W, order, steps, rng, batch=64, claims=2048,
                    min_winners=2):
    E = grid_edges(H, W)
    n = H * W
    thr = n // 2
    node = [[-1, 0] for _ in range(n)]     # (parent or -1, offset to parent)
    gsz = [0] * n                          # size - 1 / NOT_ROOT / SAT

    def walk2(vnode, vg, ends, pres):
        out = []
        for p, pre in zip(ends, pres):
            if vg[p] != NOT_ROOT:
                out += [p, 0]
                continue
            v = list(pre) if pre[0] >= 0 else list(vnode[p])
            if v[0] < 0:                   # link not visible yet: stop here
                out += [p, 0]
                continue
            r, s = p, 0
            for _ in range(64):
                q = v[0]
                s += v[1]
                if vg[q] != NOT_ROOT:
                    r = q
                    break
                w = list(vnode[q])
                if w[0] < 0:
                    r = q
                    break
                node[r] = [w[0], v[1] + w[1]]      # path splitting
                vnode[r] = node[r]
                r, v = q, w
            else:
                raise AssertionError("walk did not end")
            out += [r, s]
        return out

    def resolve(a, t, g):
        for _ in range(64):
            if g != NOT_ROOT:
                break
            v = node[a]
            assert v[0] >= 0, "the merging side sees its own links"
            t += v[1]
            a = v[0]
            g = gsz[a]
        return a, t, g

    batches = [order[i:i + batch] for i in range(0, len(order), batch)]
    pre = [([-1, 0], [-1, 0])] * len(batches[0])
    in_flight = []                         # merges of the batch being merged
    walked = None
    for i in range(len(batches) + 1):
        results = None
        if i < len(batches):
            vnode = [list(x) for x in node]
            vg = list(gsz)
            for child, target, old_child, old_target in in_flight:
                u = rng.random()
                if u < 0.3:                # nothing of this merge visible
                    vnode[child] = [-1, 0]
                    vg[child], vg[target] = old_child, old_target
                elif u < 0.6:              # size words visible, link not
                    vnode[child] = [-1, 0]
            results = []
            for j, k in enumerate(batches[i]):
                results.append(walk2(vnode, vg, E[k], pre[j]) + [steps[k]])
            if i + 1 < len(batches):
                pre = [(list(vnode[E[k][0]]), list(vnode[E[k][1]]))
                       for k in batches[i + 1]]
        in_flight = []
        if walked is not None:
            st = [list(x) for x in walked]

            def merge(a1, t1, a2, t2, e, g1, g2):
                keep1 = g2 == 0 or g1 > g2
                child, target = (a2, a1) if keep1 else (a1, a2)
                in_flight.append((child, target, gsz[child], gsz[target]))
                tot = g1 + g2 + 1
                gsz[target] = SAT if tot >= thr else tot
                gsz[child] = NOT_ROOT
                node[child] = [target, t1 - e - t2 if keep1 else t2 + e - t1]

            pend = [j for j in range(len(st)) if st[j][0] != st[j][2]]
            while pend:
                g = {}
                for j in pend:
                    a1, t1, g1 = resolve(st[j][0], st[j][1], gsz[st[j][0]])
                    a2, t2, g2 = resolve(st[j][2], st[j][3], gsz[st[j][2]])
                    st[j][:4] = [a1, t1, a2, t2]
                    g[j] = (g1, g2)
                pend = [j for j in pend if st[j][0] != st[j][2]]
                if not pend:
                    break
                claim = {}
                for j in pend:
                    for a, ga in ((st[j][0], g[j][0]), (st[j][2], g[j][1])):
                        if ga != SAT:
                            h = a % claims
                            claim[h] = min(claim.get(h, len(st)), j)
                wins = [j for j in pend
                        if all(ga == SAT or claim[a % claims] == j
                               for a, ga in ((st[j][0], g[j][0]),
                                             (st[j][2], g[j][1])))]
                assert pend[0] in wins
                if len(wins) < min_winners:
                    for j in pend:             # in edge order
                        a1, t1, g1 = resolve(st[j][0], st[j][1], gsz[st[j][0]])
                        a2, t2, g2 = resolve(st[j][2], st[j][3], gsz[st[j][2]])
                        if a1 != a2:
                            merge(a1, t1, a2, t2, st[j][4], g1, g2)
                    break
                for j in wins:
                    merge(*st[j][:5], *g[j])
                pend = [j for j in pend if j not in wins]
        walked = results
    inc = []
    for p in range(n):
        s = 0
        while gsz[p] == NOT_ROOT:
            s += node[p][1]
            p = node[p][0]
        inc.append(s)
    return np.array(inc)


@pytest.mark.parametrize("seed", range(12))
def test_merge_schedule_equals_sequential_rules(seed):
    rng = np.random.default_rng(seed)
    H, W = int(rng.integers(2, 30)), int(rng.integers(3, 30))
    E = grid_edges(H, W)
    order = list(rng.permutation(len(E)))
    steps = [int(s) for s in rng.integers(-1, 2, len(E))]
    ref = P.merge_groups(H * W, [(E[k][0], E[k][1], steps[k]) for k in order])
    for batch, claims, min_winners in ((64, 2048, 2), (8, 4, 1), (3, 16, 64)):
        got = kernel_schedule(H, W, order, steps, rng, batch, claims,
                              min_winners)
        assert np.array_equal(got, ref), (H, W, batch, claims, min_winners)


def test_merge_schedule_majority_rounding_odd_and_even_sizes():
    """The majority code is set from size - 1 >= n // 2 (more than half the
    pixels for odd and even n) and must not change a comparison."""
    rng = np.random.default_rng(99)
    for H, W in ((3, 3), (5, 7), (4, 6), (16, 16)):
        E = grid_edges(H, W)
        for _ in range(4):
            order = list(rng.permutation(len(E)))
            steps = [int(s) for s in rng.integers(-1, 2, len(E))]
            ref = P.merge_groups(H * W, [(E[k][0], E[k][1], steps[k])
                                         for k in order])
            got = kernel_schedule(H, W, order, steps, rng, 64, 2048, 1)
            assert np.array_equal(got, ref), (H, W)
