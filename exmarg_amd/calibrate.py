"""Copy 1 GiB with the library's plain copy kernel and print its time and bandwidth.

    python -m exmarg_amd.calibrate [bytes]

Run under a profiler's memory-traffic counters (FETCH_SIZE, WRITE_SIZE) this gives the counter
reading of a kernel that reads and writes exactly `bytes` each, the scale for the readings of
other kernels.
"""

import ctypes
import json
import sys

from . import _hip


def main():
    nbytes = int(sys.argv[1]) if len(sys.argv) > 1 else 1 << 30
    _hip.init(0)
    ms = ctypes.c_double()
    _hip.call("exm_copy_calibration", nbytes, ctypes.byref(ms))       # first touch
    _hip.call("exm_copy_calibration", nbytes, ctypes.byref(ms))
    print(json.dumps({"kernel": "k_copy_calibration", "bytes_read": nbytes, "bytes_written": nbytes,
                      "ms": ms.value, "GB_per_s": 2 * nbytes / ms.value / 1e6}), flush=True)


if __name__ == "__main__":
    main()
