"""Oracle for the LUE flow-routing path: a CPU restatement of LUE's algorithm, used as the checker.

Test infrastructure only (tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg)."""
