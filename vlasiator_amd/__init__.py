"""MI355X-native velocity-space acceleration (the ACC step of fmihpc/vlasiator).

The product is the HIP library vlasiator_amd/lib/libvacc.so with the C ABI of
include/vacc.h; vlasiator_amd.vacc binds it for Python callers, vlasiator_amd.host
holds the C++ host adapter, vlasiator_amd.workloads the synthetic benchmark cells.
"""
__all__ = ["vacc", "workloads"]
