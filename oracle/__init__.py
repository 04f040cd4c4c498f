"""CPU oracle for the FleCSPH per-step pipeline.  TEST INFRASTRUCTURE ONLY:
imported by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg,
never by the flecsph_amd product package."""
