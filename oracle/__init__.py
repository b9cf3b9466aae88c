"""CPU oracle (FP64 restatement of the reference's CPU algorithms).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the shearcuda_amd package.
"""
