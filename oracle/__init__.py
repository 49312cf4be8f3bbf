"""CPU oracle of the FLIP step -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may import
this package, and only as the checker or the timed CPU baseline; the product
(aquarius_amd, libaqflip) never loads it.
"""
