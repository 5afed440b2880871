"""Known-size copy for calibrating rocprofv3 memory counters.

Runs `--copies` device-to-device hipMemcpy of `--mib` MiB each (the HIP
runtime's blit kernel reads and writes exactly that many bytes), so that
FETCH_SIZE / WRITE_SIZE of that kernel, divided by the size, gives the factor
by which those counters over- or under-report HBM bytes on this device.

    rocprofv3 --pmc FETCH_SIZE -- python3 -m tgtp_amd.calib --mib 1024
"""
from __future__ import annotations

import argparse
import ctypes as C


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--mib", type=int, default=1024)
    ap.add_argument("--copies", type=int, default=3)
    a = ap.parse_args()
    hip = C.CDLL("libamdhip64.so")
    nbytes = a.mib << 20
    src, dst = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(src), C.c_size_t(nbytes)) == 0
    assert hip.hipMalloc(C.byref(dst), C.c_size_t(nbytes)) == 0
    assert hip.hipMemset(src, 1, C.c_size_t(nbytes)) == 0
    for _ in range(a.copies):
        assert hip.hipMemcpy(dst, src, C.c_size_t(nbytes), 3) == 0  # DeviceToDevice
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(src)
    hip.hipFree(dst)
    print(f"copied {a.copies} x {nbytes} bytes", flush=True)
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
