"""hipEvent timing through the HIP runtime (ctypes), for kernels launched on a
library-owned stream."""
from __future__ import annotations

import ctypes as C
import os

_hip = None


def hip() -> C.CDLL:
    global _hip
    if _hip is None:
        for p in ("/opt/rocm/lib/libamdhip64.so", "libamdhip64.so"):
            try:
                _hip = C.CDLL(p)
                break
            except OSError:
                continue
        if _hip is None:
            raise RuntimeError("libamdhip64.so not found")
        _hip.hipEventCreate.argtypes = [C.POINTER(C.c_void_p)]
        _hip.hipEventRecord.argtypes = [C.c_void_p, C.c_void_p]
        _hip.hipEventSynchronize.argtypes = [C.c_void_p]
        _hip.hipEventElapsedTime.argtypes = [C.POINTER(C.c_float), C.c_void_p, C.c_void_p]
        _hip.hipEventDestroy.argtypes = [C.c_void_p]
        _hip.hipStreamBeginCapture.argtypes = [C.c_void_p, C.c_int]
        _hip.hipStreamEndCapture.argtypes = [C.c_void_p, C.POINTER(C.c_void_p)]
        _hip.hipGraphInstantiate.argtypes = [C.POINTER(C.c_void_p), C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t]
        _hip.hipGraphLaunch.argtypes = [C.c_void_p, C.c_void_p]
        _hip.hipGraphExecDestroy.argtypes = [C.c_void_p]
        _hip.hipGraphDestroy.argtypes = [C.c_void_p]
    return _hip


class EventTimer:
    """ms between two events recorded on `stream`"""

    def __init__(self, stream):
        self.stream = stream
        self.e0, self.e1 = C.c_void_p(), C.c_void_p()
        assert hip().hipEventCreate(C.byref(self.e0)) == 0
        assert hip().hipEventCreate(C.byref(self.e1)) == 0

    def start(self):
        assert hip().hipEventRecord(self.e0, self.stream) == 0

    def stop(self) -> float:
        assert hip().hipEventRecord(self.e1, self.stream) == 0
        assert hip().hipEventSynchronize(self.e1) == 0
        ms = C.c_float()
        assert hip().hipEventElapsedTime(C.byref(ms), self.e0, self.e1) == 0
        return float(ms.value)

    def close(self):
        hip().hipEventDestroy(self.e0)
        hip().hipEventDestroy(self.e1)


def available() -> bool:
    try:
        hip()
        return os.path.exists("/dev/kfd")
    except Exception:
        return False


def graph_ms_per_call(stream, fn, reps: int, launches: int = 3) -> float:
    """fn() launches work on `stream`; it is captured `reps` times into one HIP
    graph (relaxed capture mode), which is replayed once untimed and then
    `launches` times between events: the average ms per fn() as the GPU runs
    it back to back, without the host's per-launch gaps."""
    h = hip()
    assert h.hipStreamBeginCapture(stream, 2) == 0
    try:
        for _ in range(reps):
            fn()
    finally:
        graph = C.c_void_p()
        rc = h.hipStreamEndCapture(stream, C.byref(graph))
    assert rc == 0, "stream capture failed"
    exe = C.c_void_p()
    assert h.hipGraphInstantiate(C.byref(exe), graph, None, None, 0) == 0
    try:
        assert h.hipGraphLaunch(exe, stream) == 0
        ev = EventTimer(stream)
        ev.start()
        for _ in range(launches):
            assert h.hipGraphLaunch(exe, stream) == 0
        ms = ev.stop()
        ev.close()
    finally:
        h.hipGraphExecDestroy(exe)
        h.hipGraphDestroy(graph)
    return ms / (launches * reps)
