"""CPU oracle for the HIP SPOCK solver.  Test infrastructure: imported only by tests/,
``__graft_entry__.smoke()`` and ``bench.py``'s cpu_baseline leg."""
