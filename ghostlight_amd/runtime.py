"""Device runtime handles: render contexts (one HIP stream each) and
device-resident lenses.  Thin owners of the C ABI's opaque pointers."""
from __future__ import annotations

import ctypes as C

from . import _abi


def lib():
    return _abi.load_library()


def device_available() -> bool:
    """True when a gfx950 device is usable (replaces `_cuda_available`,
    bind_renderers.cpp:518-521)."""
    try:
        return bool(lib().gl_device_available())
    except (ImportError, OSError):
        return False


def device_memory():
    """(free, total) bytes of the current device."""
    f, t = C.c_uint64(0), C.c_uint64(0)
    _abi.check(lib().gl_device_memory(C.byref(f), C.byref(t)), "gl_device_memory")
    return int(f.value), int(t.value)


class Context:
    """One HIP stream and its grow-only device buffers.  Renders on different
    contexts may run concurrently on the device."""

    def __init__(self):
        self._h = C.c_void_p()
        _abi.check(lib().gl_ctx_create(C.byref(self._h)), "gl_ctx_create")

    @property
    def handle(self):
        return self._h

    def synchronize(self):
        _abi.check(lib().gl_ctx_synchronize(self._h), "gl_ctx_synchronize")

    def last_kernel_ms(self) -> float:
        return float(lib().gl_ctx_last_kernel_ms(self._h))

    def last_splat_ms(self) -> float:
        return float(lib().gl_ctx_last_splat_ms(self._h))

    def last_splat_counts(self) -> dict:
        """Deposit-record counts of the last ghost render on this context."""
        a = (C.c_uint64 * 4)()
        _abi.check(lib().gl_ctx_last_splat_counts(self._h, a), "gl_ctx_last_splat_counts")
        return {"records": int(a[0]), "binned": int(a[1]), "chunks": int(a[2]), "staged_tiles": int(a[3])}

    def last_schedule(self) -> dict:
        """Schedule the last ghost render on this context used."""
        a = (C.c_int32 * 8)()
        _abi.check(lib().gl_ctx_last_schedule(self._h, a), "gl_ctx_last_schedule")
        return {"lambda_group": a[0], "workers": a[1], "longest_first": bool(a[2]), "candidate": a[3],
                "tuning": bool(a[4]), "work_items": a[5], "graph": bool(a[6]), "candidates": a[7]}

    def memory_bytes(self) -> int:
        """Device bytes this context's buffer cache holds."""
        return int(lib().gl_ctx_memory_bytes(self._h))

    def trim(self) -> None:
        """Release the context's cached device buffers."""
        _abi.check(lib().gl_ctx_trim(self._h), "gl_ctx_trim")

    def close(self):
        if self._h:
            lib().gl_ctx_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class DeviceLens:
    """A prescription packed for the device (copied to HBM on first use)."""

    def __init__(self, lens):
        self.surfaces, self.arena = lens.pack()
        self.n = lens.num_surfaces()
        self._h = C.c_void_p()
        _abi.check(lib().gl_lens_create(self.surfaces, self.n, _abi.fptr(self.arena), len(self.arena),
                                        C.byref(self._h)), "gl_lens_create")

    @property
    def handle(self):
        return self._h

    def __del__(self):
        try:
            if self._h:
                lib().gl_lens_destroy(self._h)
        except Exception:
            pass


def ctx_handle(ctx):
    return None if ctx is None else (ctx.handle if isinstance(ctx, Context) else ctx)
