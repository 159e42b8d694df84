"""The oracle: a CPU restatement of the reference SUQA engine.

TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg may import this package; the product package
(suqa_amd) never does.
"""
