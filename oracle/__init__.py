"""CPU oracle of the visualTur hot path -- test infrastructure only.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use
this package, as the checker or as the timed CPU baseline.  The shipped
library never imports it.
"""
