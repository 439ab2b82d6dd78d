"""eval::recall_at_k (src/eval.rs:25-47)."""


def recall_at_k(truth, found, k):
    """Mean recall@k of `found` against exact `truth` (lists of Neighbor lists)."""
    if len(truth) != len(found):
        raise ValueError("truth and found must have one entry per query")
    hits = expected = 0
    for t, f in zip(truth, found):
        tk = [n.key for n in t[:k]]
        expected += len(tk)
        hits += sum(1 for n in f[:k] if n.key in tk)
    return 1.0 if expected == 0 else hits / expected
