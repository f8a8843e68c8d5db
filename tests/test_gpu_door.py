"""The front door (vtg_door_*, src/card/session_track.c:76-139, 188-222) over the serving loop, as the reference's host
test drives it (src/host/vt_run.c:66-84, 120-148): two cameras publish frames through a descriptor ring into host
slots; one malformed message of each kind is published among them (exactly one thing wrong with each); the door counts
each as an exception (last_exception the last kind), frees its slot, and answers every good frame in order with the LK
records a direct server run of the same frames gives, bit for bit. A LIVE frame skips the frames taken before it;
the door takes no more than dev_slots unanswered frames."""
import ctypes as C

import numpy as np
import pytest

import vtgpu
from gpu_util import say
from workloads import StreamFrames

pytestmark = pytest.mark.gpu

W, H, L = 320, 240, 3


class Host:
    """The host side of the wire: a descriptor ring, frame slots and the publication count."""

    def __init__(self, nslots):
        self.desc = (vtgpu.Msg * vtgpu.DESC_RING)()
        self.nslots = nslots
        self.slots = np.zeros(nslots * W * H, np.uint8)
        self.pub = 0

    def publish(self, cam, instance, frame, priority=vtgpu.PRIO_ORDERED, bad=None):
        slot = self.pub % self.nslots
        self.slots[slot * W * H:(slot + 1) * W * H] = frame.ravel()
        m = vtgpu.Msg()
        m.magic, m.type, m.dest, m.priority = vtgpu.MSG_MAGIC, vtgpu.MSG_FRAME, vtgpu.PIPE_ALL, priority
        m.length, m.instance, m.slot, m.camera, m.acq_ns = W * H, instance, slot, cam, 1000 + instance
        if bad == vtgpu.EXC_BAD_TYPE:
            m.type = 99
        elif bad == vtgpu.EXC_BAD_SLOT:
            m.slot = 9999
        elif bad == vtgpu.EXC_BAD_LENGTH:
            m.length = 64
        m.check = vtgpu.msg_check(m)
        if bad == vtgpu.EXC_BAD_CHECK:
            m.check ^= 1
        self.desc[self.pub % vtgpu.DESC_RING] = m
        self.pub += 1


def _door(srv, sf, dev_slots, nslots=8):
    cfg = vtgpu.DoorCfg()
    cfg.cameras = 2
    for k in range(2):
        cfg.w[k], cfg.h[k], cfg.levels[k], cfg.npoints[k] = W, H, L, len(sf[k].xy) // 2
    cfg.max_points, cfg.slots, cfg.slot_bytes, cfg.dev_slots = 400, nslots, W * H, dev_slots
    return vtgpu.Door(srv, cfg, [sf[0].xy, sf[1].xy])


def test_door_exceptions_and_answers():
    sf = [StreamFrames(vtgpu, W, H, 60 + k, k=6, nx=20, ny=15) for k in range(2)]
    n = len(sf[0].xy) // 2
    srv = vtgpu.Server(1 << 30, lanes=2, max_batch=8, cfg=vtgpu.lk_cfg(15, 15))
    door = _door(srv, sf, dev_slots=4)
    host = Host(8)
    bad_at = {3: vtgpu.EXC_BAD_TYPE, 5: vtgpu.EXC_BAD_SLOT, 7: vtgpu.EXC_BAD_LENGTH, 9: vtgpu.EXC_INSTANCE_ORDER,
              11: vtgpu.EXC_BAD_CHECK}
    sent, answers = [], []
    for i in range(12):
        for cam in range(2):
            if cam == 0 and i in bad_at:
                kind = bad_at[i]
                host.publish(0, 0 if kind == vtgpu.EXC_INSTANCE_ORDER else 100 + i, sf[0][i], bad=kind)
            host.publish(cam, i + 1, sf[cam][i])
            sent.append((cam, i))
            # the host may reuse slots only once they are taken; answer what is done to make room
            while True:
                st = door.ingest(host.desc, host.pub, host.slots)
                if st["frames_taken"] == host.pub:
                    break
                r = door.result(n, wait=True)
                assert r is not None
                answers.append(r)
    while True:
        r = door.result(n, wait=True)
        if r is None:
            break
        answers.append(r)
    say("door", st)
    assert st["exceptions"] == 5 and st["last_exception"] == vtgpu.EXC_BAD_CHECK and st["skipped"] == 0
    assert [(int(a.camera), int(a.instance)) for a, _ in answers] == [(c, i + 1) for c, i in sent]
    # the same frames straight into a server give the same records
    direct = vtgpu.Server(1 << 30, lanes=2, max_batch=8, cfg=vtgpu.lk_cfg(15, 15))
    ids = [direct.open(W, H, L, 400) for _ in range(2)]
    for (a, recs), (cam, i) in zip(answers, sent):
        t = direct.submit(ids[cam], sf[cam][i], sf[cam].xy)
        ref = direct.wait(t)[0]
        assert recs.tobytes() == ref.tobytes(), (cam, i)
        assert len(recs) == (0 if i == 0 else n)
    direct.destroy()
    door.destroy()
    srv.destroy()


def test_door_live_and_backpressure():
    sf = [StreamFrames(vtgpu, W, H, 70 + k, k=6, nx=20, ny=15) for k in range(2)]
    n = len(sf[0].xy) // 2
    srv = vtgpu.Server(1 << 30, lanes=2, max_batch=8, cfg=vtgpu.lk_cfg(15, 15))
    # backpressure: two unanswered frames at most
    door = _door(srv, sf, dev_slots=2)
    host = Host(8)
    for i in range(5):
        host.publish(0, i + 1, sf[0][i])
    st = door.ingest(host.desc, host.pub, host.slots)
    assert st["frames_taken"] == 2 and st["in_flight"] == 2
    assert door.ingest(host.desc, host.pub, host.slots)["frames_taken"] == 2     # full: nothing more taken
    a, _ = door.result(n, wait=True)
    assert a.instance == 1
    st = door.ingest(host.desc, host.pub, host.slots)
    assert st["frames_taken"] == 3 and st["in_flight"] == 2
    while door.result(n, wait=True) is not None:
        pass
    st = door.ingest(host.desc, host.pub, host.slots)
    assert st["frames_taken"] == 5
    while door.result(n, wait=True) is not None:
        pass
    door.destroy()
    # LIVE: four ordered frames then a live one, all taken in one call: only the live one runs
    door = _door(srv, sf, dev_slots=8)
    host = Host(8)
    for i in range(4):
        host.publish(1, i + 1, sf[1][i])
    host.publish(1, 5, sf[1][4], priority=vtgpu.PRIO_LIVE)
    st = door.ingest(host.desc, host.pub, host.slots)
    assert st["frames_taken"] == 5 and st["skipped"] == 4 and st["submitted"] == 1
    a, recs = door.result(n, wait=True)
    assert (a.camera, a.instance, len(recs)) == (1, 5, 0)   # the camera's first frame: nothing to track from
    assert door.result(n, wait=False) is None
    door.destroy()
    srv.destroy()


def _msg(cam, instance, slot, length, priority=vtgpu.PRIO_ORDERED):
    m = vtgpu.Msg()
    m.magic, m.type, m.dest, m.priority = vtgpu.MSG_MAGIC, vtgpu.MSG_FRAME, vtgpu.PIPE_ALL, priority
    m.length, m.instance, m.slot, m.camera, m.acq_ns = length, instance, slot, cam, 1000 + instance
    m.check = vtgpu.msg_check(m)
    return m


@pytest.mark.parametrize("stereo", [0, 1])
def test_door_live_across_calls_and_slot_length(stereo):
    """A 350 x 250 camera (w x h = 87,500, not a multiple of 64): a descriptor of length w x h is BAD_LENGTH and one of
    vt_slot_bytes (rounded up to 64, twice w x h first for a stereo camera 0) is taken, as the reference's ingest
    checks it (session_track.c:82, 107). With one frame at a time at the server, a LIVE frame published in a later
    ingest call than an ordered frame still waiting skips it (build_one, session_track.c:197-206): the answers are the
    first frame and the LIVE one, the waiting frame is counted as skipped."""
    w, h = 350, 250
    one = w * h
    sb = ((2 if stereo else 1) * one + 63) & ~63
    sf = StreamFrames(vtgpu, w, h, 90, k=6, nx=12, ny=10)
    n = len(sf.xy) // 2
    srv = vtgpu.Server(1 << 30, lanes=2, max_batch=8, cfg=vtgpu.lk_cfg(15, 15))
    cfg = vtgpu.DoorCfg()
    cfg.cameras, cfg.w[0], cfg.h[0], cfg.levels[0], cfg.npoints[0] = 1, w, h, L, n
    cfg.max_points, cfg.slots, cfg.slot_bytes, cfg.dev_slots = 400, 8, sb, 8
    cfg.stereo, cfg.max_inflight = stereo, 1
    door = vtgpu.Door(srv, cfg, [sf.xy])
    desc = (vtgpu.Msg * vtgpu.DESC_RING)()
    slots = np.zeros(8 * sb, np.uint8)
    pub = [0]

    def publish(instance, frame, length=sb, priority=vtgpu.PRIO_ORDERED):
        slot = pub[0] % 8
        slots[slot * sb:slot * sb + one] = frame.ravel()
        if stereo:
            slots[slot * sb + one:slot * sb + 2 * one] = 255 - frame.ravel()
        desc[pub[0] % vtgpu.DESC_RING] = _msg(0, instance, slot, length, priority)
        pub[0] += 1

    say("door live across calls, stereo", stereo)
    publish(1, sf[0])
    st = door.ingest(desc, pub[0], slots)
    assert st["submitted"] == 1
    publish(2, sf[1])
    st = door.ingest(desc, pub[0], slots)
    assert st["submitted"] == 1 and st["in_flight"] == 2   # frame 2 waits landed behind frame 1
    publish(3, sf[2], length=one)                             # the unrounded size: dropped
    publish(4, sf[3], priority=vtgpu.PRIO_LIVE)
    st = door.ingest(desc, pub[0], slots)
    assert st["exceptions"] == 1 and st["last_exception"] == vtgpu.EXC_BAD_LENGTH and st["frames_taken"] == 4
    answers = []
    while True:
        r = door.result(n, wait=True)
        if r is None:
            break
        answers.append(r)
    st = door.ingest(desc, pub[0], slots)
    assert [int(a.instance) for a, _ in answers] == [1, 4] and st["skipped"] == 1 and st["in_flight"] == 0
    # frame 4 tracked from frame 1, the last one the stream built: the same records as a direct server run
    direct = vtgpu.Server(1 << 30, lanes=2, max_batch=8, cfg=vtgpu.lk_cfg(15, 15))
    sid = direct.open(w, h, L, 400)
    direct.wait(direct.submit(sid, sf[0], sf.xy))
    ref = direct.wait(direct.submit(sid, sf[3], sf.xy))[0]
    assert answers[1][1].tobytes() == ref.tobytes()
    direct.destroy()
    door.destroy()
    srv.destroy()
