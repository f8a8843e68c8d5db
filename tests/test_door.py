"""The front door's message check (vtg_msg_check, src/common/vt_proto.h:181-191) equals the reference's own inline
vt_msg_check (exported from its header by oracle/_ref/libvtproto.so) on random descriptors, and struct vtg_msg has
struct vt_msg's size; a flipped bit anywhere before the check word changes the check (CPU only)."""
import ctypes as C
import os

import numpy as np
import pytest

import vtgpu
from oracle_ffi import REPO

PROTO_SO = os.path.join(REPO, "oracle", "_ref", "libvtproto.so")


def _random_msgs(n, seed):
    rng = np.random.default_rng(seed)
    raw = rng.integers(0, 256, (n, C.sizeof(vtgpu.Msg)), dtype=np.uint8)
    return [vtgpu.Msg.from_buffer_copy(raw[i].tobytes()) for i in range(n)]


@pytest.mark.skipif(not os.path.exists(PROTO_SO), reason="oracle/_ref not built (no reference tree here)")
def test_msg_check_equals_reference():
    ref = C.CDLL(PROTO_SO)
    ref.ref_vt_msg_check.restype = C.c_uint64
    ref.ref_vt_msg_check.argtypes = [C.c_void_p]
    ref.ref_vt_msg_size.restype = C.c_uint
    assert ref.ref_vt_msg_size() == C.sizeof(vtgpu.Msg) == 88
    for m in _random_msgs(2000, 1):
        assert vtgpu.msg_check(m) == ref.ref_vt_msg_check(C.addressof(m))


def test_msg_check_sees_every_word():
    for m in _random_msgs(50, 2):
        c0 = vtgpu.msg_check(m)
        raw = bytearray(bytes(m))
        for byte in range(0, 80, 7):
            b = bytearray(raw)
            b[byte] ^= 0x10
            assert vtgpu.msg_check(vtgpu.Msg.from_buffer_copy(bytes(b))) != c0
        b = bytearray(raw)
        b[85] ^= 1      # the check word itself is not covered
        assert vtgpu.msg_check(vtgpu.Msg.from_buffer_copy(bytes(b))) == c0
