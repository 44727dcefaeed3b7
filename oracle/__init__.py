"""Test infrastructure: the CPU oracle of the relational hot path.

`ref_ops` restates the reference's operator algorithms on torch CPU tensors (the parity oracle
and the CPU baseline); `closed_form` states the same results in plain numpy / brute force and
pins `ref_ops`.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use this
package; the product (tqp-inmem_amd/) never imports it.
"""
