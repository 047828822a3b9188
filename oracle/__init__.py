"""CPU oracle of the H-matrix hot path (test infrastructure).

Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg import this
package, and only as the checker (or the timed CPU restatement). The product
(htool_amd) never imports it.
"""
