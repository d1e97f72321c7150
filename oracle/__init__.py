"""FP64 CPU oracle of the reference's CPU physics path — TEST INFRASTRUCTURE.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load
this package, and only as the checker / timed CPU baseline.  The product
(physsim_amd/, libgeon_hip.so) never imports, links or calls it.
"""
