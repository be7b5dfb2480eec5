"""The output cadence of a device-driven run (reference solver/run.py:347-400).

The reference writes a baseline frame at t = 0, then at every output tick records
the ledger and appends ``h``, ``u``, ``v`` (and the bed change) to the store, all
synchronously.  :class:`OutputPipeline` keeps the same frames and the same order
but takes them off the critical path: the main thread captures a frame on the
device after the steps that reach the output time and goes straight on to the next
segment, while a writer thread waits for each frame's copy and appends it to the
store.  Slots bound the frames in flight; the main thread waits only when all are
held.
"""

from __future__ import annotations

import queue
import threading

from rbs.io.frames import FrameQueue


class OutputPipeline:
    """Device frames of ``state`` appended to ``writer`` by a writer thread."""

    def __init__(self, state, writer, *, bed_change: bool = False, slots: int = 3):
        fields = ("depth", "u", "v") + (("bed_change",) if bed_change else ())
        self.frames = FrameQueue(state, fields, slots)
        self.writer = writer
        self._todo: queue.Queue = queue.Queue()
        self._free = threading.Semaphore(slots)
        self._error: BaseException | None = None
        self.captured = 0
        self.written = 0
        self._thread = threading.Thread(target=self._drain, name="rbs-output", daemon=True)
        self._thread.start()

    def capture(self) -> None:
        """Frame the state as it stands after its enqueued steps (returns at once
        unless every slot is still being written)."""
        self._free.acquire()
        if self._error is not None:
            self._free.release()
            raise RuntimeError("the output writer failed") from self._error
        slot = self.frames.capture()
        self.captured += 1
        self._todo.put(slot)

    def _drain(self) -> None:
        while True:
            slot = self._todo.get()
            if slot is None:
                return
            try:
                if self._error is None:
                    f = self.frames.wait(slot)
                    self.writer.append(f.time, f.depth, f.u, f.v, f.bed_change)
                    self.written += 1
            except BaseException as e:  # surfaced on the next capture or at close
                self._error = e
            finally:
                self.frames.release(slot)
                self._free.release()

    def close(self) -> None:
        """Write every captured frame, then free the slots."""
        if self._thread.is_alive():
            self._todo.put(None)
            self._thread.join()
        self.frames.close()
        if self._error is not None:
            raise RuntimeError("the output writer failed") from self._error


def run_with_output(run, sched, pipeline: OutputPipeline, *, ledger=None, on_output=None) -> int:
    """Drive ``run`` through ``sched``'s segments with frames at the output cadence
    (run.py:347-400): the baseline frame, then slow processes and the ledger record
    at each sync point, then the frame.  Returns the steps taken."""
    pipeline.capture()  # the frame at t = 0 (run.py:347-349)
    steps = 0
    for seg in sched.segments():
        steps += run.run_until(seg.t_stop)
        for proc, elapsed in seg.due:
            proc.advance(seg.t_stop, elapsed)
        if seg.is_output:
            if ledger is not None:
                ledger.record(run.state, seg.t_stop)
            pipeline.capture()
            if on_output is not None:
                on_output(seg.t_stop)
    return steps
