"""The seeded workload generators are deterministic and match their stated shapes."""
import numpy as np

from cuda_replace_amd import workloads


def test_uneven_documents_deterministic_and_shaped():
    a_docs, a_pairs = workloads.uneven_documents(num_docs=96, median=8 << 10)
    b_docs, b_pairs = workloads.uneven_documents(num_docs=96, median=8 << 10)
    assert a_docs == b_docs and a_pairs == b_pairs
    sizes = np.array([len(d) for d in a_docs])
    assert sizes.min() >= 1 << 10 and sizes.max() <= 32 << 20
    assert 1 <= min(len(p) for p in a_pairs) and max(len(p) for p in a_pairs) <= 16
    for prs in a_pairs:
        for pat, rep in prs:
            assert 1 <= len(pat) <= 16 and len(rep) <= 32
    # the documents are consecutive slices of one text
    text = workloads.zipf_text(int(sizes.sum()))
    assert b"".join(a_docs) == text


def test_pair_pool_patterns_occur():
    sample = workloads.zipf_text(1 << 20)
    pool = workloads.pair_pool(sample, 64, seed=4)
    assert all(sample.count(p) >= 1 for p, _ in pool)
    assert any(len(r) > len(p) for p, r in pool) and any(len(r) < len(p) for p, r in pool)
