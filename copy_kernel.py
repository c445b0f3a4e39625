"""A plain device-to-device copy of a known size, for calibrating HBM byte
counters (rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE): the HIP runtime performs a
hipMemcpy between two device buffers with its blit kernel, which reads and
writes exactly the buffer size.

python copy_kernel.py [--mib 1024] [--reps 5]
"""
from __future__ import annotations

import argparse
import ctypes as C
import json


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--mib", type=int, default=1024)
    ap.add_argument("--reps", type=int, default=5)
    a = ap.parse_args(argv)
    hip = C.CDLL("libamdhip64.so")
    n = a.mib << 20
    src, dst = C.c_void_p(), C.c_void_p()
    for p in (src, dst):
        if hip.hipMalloc(C.byref(p), C.c_size_t(n)) != 0:
            raise SystemExit("hipMalloc failed")
    if hip.hipMemset(src, 1, C.c_size_t(n)) != 0:
        raise SystemExit("hipMemset failed")
    hipMemcpyDeviceToDevice = 3
    for _ in range(a.reps):
        if hip.hipMemcpy(dst, src, C.c_size_t(n), hipMemcpyDeviceToDevice) != 0:
            raise SystemExit("hipMemcpy failed")
    hip.hipDeviceSynchronize()
    hip.hipFree(src)
    hip.hipFree(dst)
    print(json.dumps({"copy_bytes": n, "reps": a.reps}))


if __name__ == "__main__":
    main()
