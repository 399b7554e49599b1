"""Pure-Python restatement of the VDIF frame sequencer, for small inputs.

TEST INFRASTRUCTURE ONLY (see oracle.py).  Follows PIPELINE
(vdifFILE2dedisp_PIPELINE.cu) line by line in behaviour:
  initframePOL0/1     :829-883   header + payload read
  checkframeskip0/1   :888-930   gap -> zero frames, real frame parked
  readoneframePOL0/1  :934-1025
  readtwoframe        :1031-1040
  initskip01          :1044-1052  zero payload for gaps
  initframe01         :1058-1131  thread alignment, cal-cycle start
  filltimebufferHOST  :1135-1171  NLOOP-byte bundles with partial carry
plus a builder of synthetic VDIF byte streams (header layout vdifio.h:53-72).
"""
from __future__ import annotations

import struct

FRAMELEN, HDRLEN, DATALEN, FPS = 5032, 32, 5000, 25600


def make_header(seconds=0, frame=0, thread=0, epoch=30, invalid=0, framelength8=629, nbits=7,
                station=0x5856, legacy=0, nchan=0, version=0, iscomplex=0) -> bytes:
    w0 = (seconds & 0x3FFFFFFF) | ((legacy & 1) << 30) | ((invalid & 1) << 31)
    w1 = (frame & 0xFFFFFF) | ((epoch & 0x3F) << 24)
    w2 = (framelength8 & 0xFFFFFF) | ((nchan & 0x1F) << 24) | ((version & 7) << 29)
    w3 = (station & 0xFFFF) | ((thread & 0x3FF) << 16) | ((nbits & 0x1F) << 26) | ((iscomplex & 1) << 31)
    return struct.pack("<8I", w0, w1, w2, w3, 0, 0, 0, 0)


def parse_header(h: bytes) -> dict:
    w = struct.unpack("<8I", h[:32])
    return {"seconds": w[0] & 0x3FFFFFFF, "invalid": w[0] >> 31, "frame": w[1] & 0xFFFFFF,
            "epoch": (w[1] >> 24) & 0x3F, "thread": (w[3] >> 16) & 0x3FF}


def build_stream(frames) -> bytes:
    """frames: iterable of (seconds, frame, thread, payload[5000])."""
    out = bytearray()
    for sec, fn, th, pay in frames:
        assert len(pay) == DATALEN
        out += make_header(seconds=sec, frame=fn, thread=th)
        out += bytes(pay)
    return bytes(out)


class RefSequencer:
    """seconds=True selects the seconds-aware gap rule and init checks of
    TIMESERIES (vdifDADA2timeseries.cu:893-937,1095-1140)."""

    def __init__(self, data: bytes, seconds: bool = False):
        self.seconds = seconds
        self.expected_sec = [0, 0]
        self.data = data
        self.pos = 0
        self.hdr = [b"", b""]
        self.buf = [bytearray(DATALEN), bytearray(DATALEN)]
        self.tmp = [bytearray(DATALEN), bytearray(DATALEN)]
        self.skip = [bytearray(DATALEN), bytearray(DATALEN)]   # initskip01: zeros
        self.expected = [0, 0]
        self.nskip = [0, 0]
        self.npartfill = 0
        self.npartleft = 0
        self.log = []

    def _read(self, n):
        if self.pos + n > len(self.data):
            self.pos = len(self.data)
            return None
        b = self.data[self.pos:self.pos + n]
        self.pos += n
        return b

    def initframe(self, pol):
        h = self._read(HDRLEN)
        if h is None:
            return 1
        d = self._read(DATALEN)
        if d is None:
            return 1
        self.hdr[pol] = h
        self.buf[pol] = bytearray(d)
        return 0

    def fnum(self, pol):
        return parse_header(self.hdr[pol])["frame"]

    def tid(self, pol):
        return parse_header(self.hdr[pol])["thread"]

    def fsec(self, pol):
        return parse_header(self.hdr[pol])["seconds"]

    def checkframeskip(self, pol):
        f = self.fnum(pol)
        if self.seconds:
            s = self.fsec(pol)
            if f != self.expected[pol] or s != self.expected_sec[pol]:
                self.nskip[pol] = f - self.expected[pol] + (s - self.expected_sec[pol]) * FPS
                self.tmp[pol] = bytearray(self.buf[pol])
                self.buf[pol] = bytearray(self.skip[pol])
            self.expected[pol] = (f + 1) % FPS
            self.expected_sec[pol] = s + (1 if self.expected[pol] == 0 else 0)
            return
        if f != self.expected[pol]:
            self.nskip[pol] = f - self.expected[pol]
            if self.nskip[pol] < 0:
                self.nskip[pol] += FPS
            self.tmp[pol] = bytearray(self.buf[pol])
            self.buf[pol] = bytearray(self.skip[pol])
        self.expected[pol] = (f + 1) % FPS

    def readoneframe(self, pol):
        if self.nskip[pol]:
            self.nskip[pol] -= 1
            if self.nskip[pol] == 0:
                self.buf[pol] = bytearray(self.tmp[pol])
            return 0
        if self.initframe(pol):
            return 1
        self.checkframeskip(pol)
        return 0

    def readtwoframe(self):
        if self.readoneframe(0):
            return 1
        if self.readoneframe(1):
            return 1
        return 0

    def initframe01(self):
        """0 ok, 1 = end of input, 2 = alignment failure."""
        if self.initframe(0):
            return 1
        if self.tid(0) != 0:
            if self.initframe(0):
                return 1
        if self.tid(0) != 0:
            return 2
        if self.initframe(1):
            return 1
        if self.tid(1) != 1:
            return 2
        if self.fnum(0) != self.fnum(1) or (self.seconds and self.fsec(0) != self.fsec(1)):
            return 2
        f0 = self.fnum(0)
        while f0 % 2560 != 0:
            if self.initframe(0):
                return 1
            if self.initframe(1):
                return 1
            f0 = self.fnum(0)
        if f0 != self.fnum(1) or (self.seconds and self.fsec(0) != self.fsec(1)):
            return 2
        self.expected = [(f0 + 1) % FPS, (self.fnum(1) + 1) % FPS]
        if self.seconds:
            self.expected_sec = [self.fsec(p) + (1 if self.expected[p] == 0 else 0) for p in (0, 1)]
        return 0

    def fill(self, nloop):
        """Returns (status, pol0 bytes, pol1 bytes); status 1 = bundle dropped."""
        t0, t1 = bytearray(), bytearray()
        if self.npartleft:
            t0 += self.buf[0][self.npartfill:self.npartfill + self.npartleft]
            t1 += self.buf[1][self.npartfill:self.npartfill + self.npartleft]
            if self.readtwoframe():
                return 1, None, None
        ntot = nloop - self.npartleft
        nwhole = ntot // DATALEN
        self.npartfill = ntot % DATALEN
        self.npartleft = DATALEN - self.npartfill
        for _ in range(nwhole):
            t0 += self.buf[0]
            t1 += self.buf[1]
            if self.readtwoframe():
                return 1, None, None
        if self.npartfill:
            t0 += self.buf[0][:self.npartfill]
            t1 += self.buf[1][:self.npartfill]
        return 0, bytes(t0), bytes(t1)


def reference_bundles(data: bytes, nloop: int, max_bundles=10**9, seconds=False):
    """All bundles the reference main loop would process, and the init status."""
    s = RefSequencer(data, seconds)
    st = s.initframe01()
    if st:
        return st, []
    out = []
    while len(out) < max_bundles:
        st, a, b = s.fill(nloop)
        if st:
            break
        out.append((a, b))
    return 0, out
