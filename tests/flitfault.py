"""Planted faults in detection-packet flits [n][L][fw] (u32 words, flit bit i
in word i // 32): the receive checks of the writeback are tested on these.
Every fault is planted on the host into a copy of the flits."""
import numpy as np


def get_bits(flit, lsb, n):
    return sum(((int(flit[(lsb + i) // 32]) >> ((lsb + i) % 32)) & 1) << i for i in range(n))


def set_bits(flit, lsb, n, v):
    for i in range(n):
        w, b = divmod(lsb + i, 32)
        x = (int(flit[w]) & ~(1 << b)) | (((v >> i) & 1) << b)
        flit[w] = np.uint32(x & 0xFFFFFFFF)


def reseal(flit, packet_w):
    """Odd parity over the flit's packet_w + 5 bits, through bit 0."""
    set_bits(flit, 0, 1, 0)
    ones = sum(get_bits(flit, i, 1) for i in range(packet_w + 5))
    set_bits(flit, 0, 1, 1 if ones % 2 == 0 else 0)


# Header fields: (lsb within the header, width); the header starts at flit bit 5.
HDR = {"dest": (0, 4), "src": (4, 5), "vc": (9, 2), "type": (11, 3), "len": (14, 6), "seq": (20, 16)}

# Fault kinds whose counts packet_model::Scoreboard reports the same way as the
# writeback (one packet-local violation each); the others re-frame the
# Scoreboard's stream.
AGREEING = ("data", "parity", "type", "hvc", "dest", "seq", "len")
ALL = AGREEING + ("sof_body", "eof_clear", "flit_vc", "hdr_eof", "hdr_sof")


def plant(flits, k, kind, packet_w):
    """A copy of flits with one fault of `kind` in packet k."""
    fl = flits.copy()
    L = fl.shape[1]
    h = fl[k, 0]

    def hdr_field(name, fn):
        lsb, n = HDR[name]
        set_bits(h, 5 + lsb, n, fn(get_bits(h, 5 + lsb, n)) & ((1 << n) - 1))
        reseal(h, packet_w)

    if kind == "data":
        f = fl[k, 1]
        set_bits(f, 5 + 3, 1, 1 - get_bits(f, 5 + 3, 1))
    elif kind == "parity":
        set_bits(h, 0, 1, 1 - get_bits(h, 0, 1))
    elif kind == "type":
        hdr_field("type", lambda t: 7)
    elif kind == "hvc":
        hdr_field("vc", lambda v: v + 1)
    elif kind == "dest":
        hdr_field("dest", lambda d: d + 1)
    elif kind == "seq":
        hdr_field("seq", lambda s: s ^ 0x8000)
    elif kind == "len":
        hdr_field("len", lambda n: n + 1)
    elif kind == "sof_body":
        set_bits(fl[k, 1], 3, 1, 1)
        reseal(fl[k, 1], packet_w)
    elif kind == "eof_clear":
        set_bits(fl[k, L - 1], 4, 1, 0)
        reseal(fl[k, L - 1], packet_w)
    elif kind == "flit_vc":
        f = fl[k, L - 1]
        set_bits(f, 1, 2, (get_bits(f, 1, 2) + 1) & 3)
        reseal(f, packet_w)
    elif kind == "hdr_eof":
        set_bits(h, 4, 1, 1)
        reseal(h, packet_w)
    elif kind == "hdr_sof":
        set_bits(h, 3, 1, 0)
        reseal(h, packet_w)
    else:
        raise ValueError(kind)
    return fl


def ring_of_events(events, ring_words=None, addr0=0):
    """The ring a clean writeback leaves: event k's low 176 bits, zero-extended
    to 512, at word (addr0 + k) mod ring_words."""
    ev = np.asarray(events, np.uint64).reshape(-1, 3)
    n = ev.shape[0]
    R = max(n, 1) if ring_words is None else ring_words
    ring = np.zeros((R, 16), np.uint32)
    w = ev.copy()
    w[:, 2] &= np.uint64((1 << 48) - 1)
    words = w.view(np.uint32).reshape(n, 6)
    for k in range(n):
        ring[(addr0 + k) % R, :6] = words[k]
    return ring
