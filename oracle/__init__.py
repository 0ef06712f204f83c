"""Oracle for the NUFFT path: CPU restatement of the reference algorithm and an
FP64 exact NUDFT.  Test infrastructure only: imported by tests/, by
__graft_entry__.smoke() and by bench.py's cpu_baseline leg, never by the
product package."""
