"""Device-to-device copy of a known size (HIP runtime blit kernel), for two uses:

* the achievable copy bandwidth of the GPU, reported beside the 8 TB/s vendor peak;
* calibration of the rocprofv3 FETCH_SIZE / WRITE_SIZE counters: the counters read on
  this copy are divided by its known byte count to scale the remap kernel's readings.

    python3 -m vlasiator_amd.hbm_probe [--gib 2] [--iters 5]
"""
import argparse
import ctypes as C
import json


def measure(gib: float = 2.0, iters: int = 5) -> dict:
    """Times iters device-to-device copies of gib GiB on the default stream (after one
    warm-up copy); returns the copy size, the average time and the read + write rate."""
    hip = C.CDLL("libamdhip64.so")
    nbytes = int(gib * (1 << 30))
    a, b = C.c_void_p(), C.c_void_p()
    if hip.hipMalloc(C.byref(a), C.c_size_t(nbytes)) != 0:
        raise RuntimeError("hipMalloc failed")
    try:
        if hip.hipMalloc(C.byref(b), C.c_size_t(nbytes)) != 0:
            raise RuntimeError("hipMalloc failed")
        try:
            assert hip.hipMemset(a, 1, C.c_size_t(nbytes)) == 0
            e0, e1 = C.c_void_p(), C.c_void_p()
            hip.hipEventCreate(C.byref(e0))
            hip.hipEventCreate(C.byref(e1))
            hipMemcpyDeviceToDevice = 3
            hip.hipMemcpy(b, a, C.c_size_t(nbytes), hipMemcpyDeviceToDevice)   # warm-up
            hip.hipDeviceSynchronize()
            hip.hipEventRecord(e0, None)
            for _ in range(iters):
                hip.hipMemcpyAsync(b, a, C.c_size_t(nbytes), hipMemcpyDeviceToDevice, None)
            hip.hipEventRecord(e1, None)
            hip.hipEventSynchronize(e1)
            ms = C.c_float()
            hip.hipEventElapsedTime(C.byref(ms), e0, e1)
            hip.hipEventDestroy(e0)
            hip.hipEventDestroy(e1)
            per = ms.value / iters
            return {"copy_bytes": nbytes, "read_plus_write_bytes": 2 * nbytes, "avg_ms": per,
                    "copy_GBps": 2 * nbytes / (per / 1e3) / 1e9}
        finally:
            hip.hipFree(b)
    finally:
        hip.hipFree(a)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gib", type=float, default=2.0)
    ap.add_argument("--iters", type=int, default=5)
    args = ap.parse_args()
    print(json.dumps(measure(args.gib, args.iters)))


if __name__ == "__main__":
    main()
