"""MI355X-native excursion-set reionization path (find_HII_bubbles, ComputeTs, delta_Tb).

The compute lives in the HIP library ``libmrx_hip.so`` behind the C ABI of
include/mrx_reion.h; this package is its host-side mirror of the reference's
interface.
"""
