"""CPU oracle of the direct Coulomb summation path.

Test infrastructure only: imported by tests/, by ``__graft_entry__.smoke()`` and by the
CPU-baseline leg of ``bench.py``, always as the checker or the timed CPU baseline, never as
part of the GPU library's computation.
"""
