"""Partitioner pins: the reference's own V3Partition self-test known answers
(src/V3Partition.cpp:1517-1621) and structural invariants."""
import random

from rtlflow_amd.partition import Graph, contract, partition, partition_stats, step_cost


def test_self_test_x_known_answers():
    # center node with 50 inputs and 50 outputs, all cost 1, cpLimit 20
    # (PartContraction::selfTestX): longest cp 19, total 101, 14 vertices, 13 edges
    costs = [1] * 101
    edges = [(i + 1, 0) for i in range(50)] + [(0, 51 + i) for i in range(50)]
    g = Graph(costs, edges)
    contract(g, 20)
    g.remove_transitive()
    alive = g.alive()
    cp = {}
    for i in g.topo():
        t = g.m[i]
        cp[i] = t.cost + max((cp[p] for p in t.preds), default=0)
    assert max(cp.values()) == 19
    assert sum(t.cost for t in alive) == 101
    assert len(alive) == 14
    assert sum(len(t.succs) for t in alive) == 13


def test_self_test_chain_merges_to_one():
    # PartContraction::selfTestChain: a 500-long chain with cp limit 1000 -> 1 vertex
    n = 500
    g = Graph([1] * n, [(i, i + 1) for i in range(n - 1)])
    contract(g, 2 * n)
    assert len(g.alive()) == 1


def test_step_cost_within_five_percent():
    for c in list(range(1, 2000)) + [10**6, 12345678]:
        s = step_cost(c)
        assert c <= s <= c * 11 // 10 + 1


def test_partition_is_a_valid_coarsening():
    rnd = random.Random(5)
    n = 400
    edges = set()
    for b in range(1, n):
        for _ in range(rnd.randint(0, 3)):
            edges.add((rnd.randrange(b), b))
    costs = [rnd.randint(1, 50) for _ in range(n)]
    for thr in (1, 2, 4, 8):
        ms = partition(costs, sorted(edges), thr, order=list(range(n)))
        owner = {}
        for k, t in enumerate(ms):
            for v in t.vids:
                assert v not in owner
                owner[v] = k
            assert t.vids == sorted(t.vids)
        assert len(owner) == n
        pos = {t.id: k for k, t in enumerate(ms)}
        for a, b in edges:
            # every vertex edge is respected by the mtask order
            assert owner[a] == owner[b] or pos[ms[owner[a]].id] < pos[ms[owner[b]].id]
        st = partition_stats(ms)
        assert st["total"] == sum(costs)
