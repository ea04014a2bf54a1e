"""The oracle: a CPU restatement of dinara-align's alignment path, for tests and the CPU baseline only.

Test infrastructure. Only tests/, ``__graft_entry__.smoke()`` and ``bench.py``'s cpu_baseline leg import
this package, and there only as the checker (or as the timed CPU baseline beside the GPU). The product
(``dinara-align_amd``) never imports or links it.
"""
