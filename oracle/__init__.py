"""CPU oracle of the split-explicit step - test infrastructure only.

Only the tests, the smoke check and the CPU-baseline leg of the benchmark
import this package, and only as the checker or timed baseline.
"""
