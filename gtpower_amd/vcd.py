"""Device VCD ingest (include/gtpower.h, gtp_vcd_*): VCD text parsed on the
GPU into the per-(id, bit) event streams the engine evaluates.

    with VcdText(device=0) as v:
        info = v.parse(open("dump.vcd", "rb").read())
        streams = {name: first for name, _id, _w, _i, first in v.vars()}
        eng.read_vcd(v, clk_period)      # HipPower: the design's nets are the streams

Mirrors VcdReader::read (power/VcdReader.cc:80-131) and the per-bit event
store build of the CUDA path (power/cuda/CudaPower.cu:1874-2086).  A missing
engine or device raises: there is no CPU fallback.
"""
import ctypes as C

import numpy as np

from . import abi


class VcdText:
    def __init__(self, device=0):
        self._lib = abi.load_engine()
        h = C.c_void_p()
        st = self._lib.gtp_vcd_create(int(device), C.byref(h))
        if st != 0:
            raise abi.GtpError(st, "gtp_vcd_create", "no such device")
        self.handle = h
        self.info = None

    def _check(self, st, where):
        if st != 0:
            raise abi.GtpError(st, where, self._lib.gtp_vcd_last_error(self.handle).decode())

    def parse(self, text):
        """Parse VCD text (bytes); returns the gtp_vcd_info fields as a dict."""
        buf = bytes(text)
        inf = abi.GtpVcdInfo()
        self._check(self._lib.gtp_vcd_parse(self.handle, buf, len(buf), C.byref(inf)), "gtp_vcd_parse")
        self.info = {k: getattr(inf, k) for k, _ in abi.GtpVcdInfo._fields_}
        return self.info

    def vars(self):
        """[(name, id, width, id index, first stream)] in declaration order."""
        out = []
        vi = abi.GtpVcdVarInfo()
        for i in range(self.info["n_var"]):
            self._check(self._lib.gtp_vcd_var(self.handle, i, C.byref(vi)), "gtp_vcd_var")
            out.append((vi.name.decode(), vi.id.decode(), vi.width, vi.id_index, vi.first_stream))
        return out

    def events(self):
        """(offsets [n_stream + 1], times [n_event], values [n_event])."""
        S, E = self.info["n_stream"], self.info["n_event"]
        off = np.zeros(S + 1, np.int64)
        t = np.zeros(max(E, 1), np.int64)
        v = np.zeros(max(E, 1), np.uint8)
        self._check(self._lib.gtp_vcd_events(self.handle, abi.ptr(off, C.c_int64), abi.ptr(t, C.c_int64),
                                             abi.ptr(v, C.c_uint8)), "gtp_vcd_events")
        return off, t[:E], v[:E]

    def close(self):
        if self.handle:
            self._lib.gtp_vcd_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()
