"""CPU oracle: test infrastructure only (tests/, smoke(), bench.py's cpu_baseline).
The product never imports it."""
