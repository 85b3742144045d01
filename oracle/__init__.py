"""CPU restatement of the reference alignment core: test infrastructure only.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this package, and only as the checker / the timed CPU baseline.  The product
package (esperanto_amd) never imports it.
"""
