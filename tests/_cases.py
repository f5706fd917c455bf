"""Mixed encode/decode cases shared by the batched- and serving-form GPU tests."""
import numpy as np

from quicktex_amd import synth
from quicktex_amd.capi import FMT_BC1, FMT_BC3, FMT_BC4, FMT_BC5, OP_DECODE, OP_ENCODE, job

SIZES = [(1, 1), (3, 5), (4, 4), (17, 9), (64, 64), (130, 66), (256, 32), (33, 200), (8, 2), (100, 100), (5, 4),
         (96, 12)]
BPB = {FMT_BC1: 8, FMT_BC3: 16, FMT_BC4: 8, FMT_BC5: 16}


def nblocks(w, h):
    return ((w + 3) // 4) * ((h + 3) // 4)


class Case:
    """One job: host input, its expected output and device buffers."""

    def __init__(self, i, rng, qtx, size=None):
        w, h = size if size is not None else SIZES[i % len(SIZES)]
        self.w, self.h = w, h
        self.fmt = (FMT_BC1, FMT_BC3, FMT_BC4, FMT_BC5)[i % 4]
        self.op = OP_DECODE if (i // 4) % 3 == 2 else OP_ENCODE
        self.interp = int(rng.integers(0, 4))
        self.level = int(rng.integers(0, 20)) if w * h <= 64 * 64 else int(rng.integers(0, 19))
        self.mode = 4 if self.fmt == FMT_BC3 else int(rng.choice([4, 3, 0x13]))
        self.ch0 = int(rng.integers(0, 4))
        self.ch1 = int(rng.integers(0, 4))
        self.write_alpha = bool(rng.integers(0, 2))
        self.params = qtx.params(self.level, self.mode, self.interp)
        nb = nblocks(w, h)
        if self.op == OP_ENCODE:
            img = synth.mixture_rgba(w, h, seed=0x5EED1000 + i, near_black=0.2)
            img[..., 3] = rng.integers(0, 256, (h, w), dtype=np.uint8) if self.fmt != FMT_BC1 else 255
            self.src = img
            self.out_bytes = nb * BPB[self.fmt]
        else:
            self.src = rng.integers(0, 256, nb * BPB[self.fmt], dtype=np.uint8)
            self.out_bytes = w * h * 4

    def cpu(self, checker, oracle):
        from tests._cpu_libs import reference_zi
        # level 1 runs the least-squares endpoints: the zero-initialising reference build defines them
        c = (reference_zi() or oracle) if self.level == 1 else checker
        if self.op == OP_ENCODE:
            if self.fmt == FMT_BC1:
                return c.bc1_encode(self.src, self.params)
            if self.fmt == FMT_BC3:
                return c.bc3_encode(self.src, self.params)
            if self.fmt == FMT_BC4:
                return c.bc4_encode(self.src, self.ch0)
            return c.bc5_encode(self.src, self.ch0, self.ch1)
        if self.fmt == FMT_BC1:
            return c.bc1_decode(self.src, self.w, self.h, self.interp, self.write_alpha)
        if self.fmt == FMT_BC3:
            return c.bc3_decode(self.src, self.w, self.h, self.interp)
        if self.fmt == FMT_BC4:
            return c.bc4_decode(self.src, self.w, self.h, self.ch0)
        return c.bc5_decode(self.src, self.w, self.h, self.ch0, self.ch1)

    def job(self, d_src, d_dst):
        return job(self.op, self.fmt, self.w, self.h, d_src, d_dst, params=self.params, chan0=self.ch0,
                   chan1=self.ch1, write_alpha=self.write_alpha, interpolator=self.interp)

    def host_plain(self, qtx):
        """The plain host entry point (one request alone): output bytes."""
        w, h = self.w, self.h
        if self.op == OP_ENCODE:
            if self.fmt == FMT_BC1:
                return qtx.bc1_encode(self.src, self.params)
            if self.fmt == FMT_BC3:
                return qtx.bc3_encode(self.src, self.params)
            if self.fmt == FMT_BC4:
                return qtx.bc4_encode(self.src, self.ch0)
            return qtx.bc5_encode(self.src, self.ch0, self.ch1)
        if self.fmt == FMT_BC1:
            return qtx.bc1_decode(self.src, w, h, self.interp, self.write_alpha).reshape(-1)
        if self.fmt == FMT_BC3:
            return qtx.bc3_decode(self.src, w, h, self.interp).reshape(-1)
        if self.fmt == FMT_BC4:
            return qtx.bc4_decode(self.src, w, h, self.ch0).reshape(-1)
        return qtx.bc5_decode(self.src, w, h, self.ch0, self.ch1).reshape(-1)

    def plain(self, qtx, d_src, d_dst, stream):
        """The plain (one texture per launch) device entry point."""
        w, h = self.w, self.h
        if self.op == OP_ENCODE:
            if self.fmt == FMT_BC1:
                qtx.bc1_encode_device(self.params, d_src, w, h, d_dst, stream)
            elif self.fmt == FMT_BC3:
                qtx.bc3_encode_device(self.params, d_src, w, h, d_dst, stream)
            elif self.fmt == FMT_BC4:
                qtx.bc4_encode_device(self.ch0, d_src, w, h, d_dst, stream)
            else:
                qtx.bc5_encode_device(self.ch0, self.ch1, d_src, w, h, d_dst, stream)
        else:
            if self.fmt == FMT_BC1:
                qtx.bc1_decode_device(self.interp, self.write_alpha, d_src, w, h, d_dst, stream)
            elif self.fmt == FMT_BC3:
                qtx.bc3_decode_device(self.interp, d_src, w, h, d_dst, stream)
            elif self.fmt == FMT_BC4:
                qtx.bc4_decode_device(self.ch0, d_src, w, h, d_dst, stream)
            else:
                qtx.bc5_decode_device(self.ch0, self.ch1, d_src, w, h, d_dst, stream)
