"""CPU oracle (test infrastructure only): a C++ restatement of the reference path.

Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg may import this
package, as the checker or as the timed CPU baseline.  The product path never does.
"""
