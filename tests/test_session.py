"""The lineage logic of axtb_amd.session (crates/axtb-py/src/engine.rs:403-470) on the CPU: which
requests run together and which moments each starts from, with the GPU batch replaced by a
recorder (no device needed)."""
import numpy as np

from axtb_amd.session import Session


def fake_session():
    s = Session([8, 1, 1])
    s.rounds = []

    def run(positions, starts, accuracy, gradient):
        tags = [float(np.asarray(x).reshape(-1)[0]) for x in positions]
        s.rounds.append([(t, None if m is None else float(m["shell_charges"][0])) for t, m in zip(tags, starts)])
        return [{"status": 0, "energy": -t, "shell_charges": np.array([t]), "dipoles": np.zeros((3, 3)),
                 "quadrupoles": np.zeros((3, 6))} for t in tags]

    s._run = run
    return s


def geo(tag):
    x = np.zeros((3, 3))
    x[0, 0] = tag
    return x


def test_lineages_seeds_and_commits():
    s = fake_session()
    reqs = [dict(positions=geo(0), lineage="a"),
            dict(positions=geo(1), lineage="a", commit=False),
            dict(positions=geo(2), lineage="b", seeds=["a"]),
            dict(positions=geo(3), lineage="a"),
            dict(positions=geo(4), lineage="c", seeds=["b", "a"]),
            dict(positions=geo(5), lineage="b")]
    out = s.compute_many(reqs)
    # a's requests in order; b waits for a's requests before it (a1 is not committed, so b0 and
    # a2 start from a0); c starts from b0 (its first seed with history), b1 from b0
    assert s.rounds == [[(0.0, None)], [(1.0, 0.0)], [(2.0, 0.0), (3.0, 0.0)], [(4.0, 2.0), (5.0, 2.0)]]
    assert [r["from"] for r in out] == [None, "a", "a", "a", "b", "b"]
    assert {k: float(v["shell_charges"][0]) for k, v in s.history.items()} == {"a": 3.0, "b": 5.0, "c": 4.0}
    # the history carries over to the next call; a new lineage without seeds starts fresh
    out = s.compute_many([dict(positions=geo(6), lineage="b"), dict(positions=geo(7), lineage="d")])
    assert s.rounds[-1] == [(6.0, 5.0), (7.0, None)]
    assert [r["from"] for r in out] == ["b", None]


def test_one_at_a_time_equals_many():
    """compute_many's starts equal those of the same requests through compute, one at a time."""
    reqs = [dict(positions=geo(t), lineage=l, seeds=sd) for t, l, sd in
            [(0, "x", []), (1, "y", ["x"]), (2, "x", []), (3, "z", ["y"]), (4, "y", [])]]
    many = fake_session()
    many.compute_many(reqs)
    one = fake_session()
    for q in reqs:
        one.compute(q["positions"], lineage=q["lineage"], seeds=q["seeds"])
    starts = lambda s: sorted(x for r in s.rounds for x in r)
    assert starts(many) == starts(one)
