"""CPU oracle (test infrastructure only): FP64 restatements of the reference's
Strang-splitting path used to check libgpe_hip.  Only tests/, the smoke()
entry point and bench.py's cpu_baseline leg may import this package."""
