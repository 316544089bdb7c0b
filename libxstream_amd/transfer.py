"""Per-call parameter-stack transfer, overlapped with the stack kernels (the DBCSR calling pattern).

DBCSR fills a host parameter stack for every libsmm_acc_process call and copies it to the device on
the call's stream (libxstream_mem.c:1094-1127 and :1320-1403; acc_bench.c:288-307 times the same
copy); the A/B/C pools stay resident. StackPipeline keeps nslots device stacks, each with its own
pinned host staging buffer, and runs the copies on an upload stream and the stack kernels on a compute
stream. Events order every slot's reuse:
  - up[s]: the copy into slot s is done; the compute stream waits for it before the launch;
  - done[s]: the launch that reads slot s is done. The host waits for it before it refills the staging
    buffer of s and enqueues the next copy into s (that launch waited for its copy, so both the staging
    buffer and the device slot are free).
So the copy of stack i+1 runs while stack i is multiplied. With nslots = 1 and one stream for both,
the same code gives the serialised order (copy, multiply, copy, ...).

The reuse wait is on the host on purpose. Making the upload stream wait for done[s] on the device
instead (a copy waiting for a kernel on another queue) measured 0.33 ms per 30000-entry stack against
0.09 ms with the host wait and 0.10 ms serialised (23x23x23 FP64, MI355X).
"""
import numpy as np

from . import DeviceBuffer, Event, HostBuffer


class StackPipeline:
    def __init__(self, max_entries, run_stream, upload_stream=None, nslots=2, record_ints=3):
        self.run_stream = run_stream
        self.upload_stream = run_stream if upload_stream is None else upload_stream
        self.nslots = int(nslots)
        self.record_ints = int(record_ints)
        self.max_entries = int(max_entries)
        n = self.record_ints * self.max_entries
        self.dev = [DeviceBuffer(4 * max(n, 1)) for _ in range(self.nslots)]
        self.host = [HostBuffer(np.int32, max(n, 1), self.upload_stream) for _ in range(self.nslots)]
        self.up = [Event() for _ in range(self.nslots)]
        self.done = [Event() for _ in range(self.nslots)]
        self.used = [False] * self.nslots
        self.count = 0

    def submit(self, stack, launch):
        """Copy stack (entries x record_ints int32) through the next slot and call launch(dev, entries)
        on the compute stream once the copy is done. Returns launch's result."""
        stack = np.ascontiguousarray(stack, dtype=np.int32).reshape(-1, self.record_ints)
        entries = stack.shape[0]
        if entries > self.max_entries:
            raise ValueError("stack of %d entries exceeds the pipeline's %d" % (entries, self.max_entries))
        s = self.count % self.nslots
        self.count += 1
        if self.used[s]:
            self.done[s].sync()  # the launch that read this slot (and so the copy into it) has finished
        self.host[s].array[:stack.size] = stack.ravel()
        self.host[s].copy_to(self.dev[s], self.upload_stream, 4 * stack.size)
        self.up[s].record(self.upload_stream)
        if self.upload_stream is not self.run_stream:
            self.run_stream.wait(self.up[s])
        result = launch(self.dev[s], entries)
        self.done[s].record(self.run_stream)
        self.used[s] = True
        return result

    def sync(self):
        self.upload_stream.sync()
        self.run_stream.sync()

    def close(self):
        self.sync()
        for x in self.dev + self.host + self.up + self.done:
            x.free() if hasattr(x, "free") else x.close()
        self.dev, self.host, self.up, self.done = [], [], [], []
