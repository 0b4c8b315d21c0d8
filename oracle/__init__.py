"""CPU oracle for the incremental match search (test infrastructure only).

Only tests, the smoke check and the benchmark's CPU-baseline leg import this
package; the engine never does.
"""
