"""Per-shape device-workspace cache of the adapter.

Same contract as the reference's auto-cache (fc/_workspace.py:42-133): key =
(T, N) per operation, one lock per key (workspaces are not thread-safe),
clear() takes each key's lock before releasing, a failing factory disables
the cache (fail-safe) until the next clear(). The C++ workspace itself keys
its buffers by device and grows them on demand, so mask on/off alternation
and device switches stay correct.
"""
import threading
from contextlib import contextmanager

_CACHES = []


def _clear_all():
    for c in _CACHES:
        c.clear()


class _WorkspaceCache:
    def __init__(self, factory):
        self._factory = factory
        self._items = {}
        self._guard = threading.Lock()
        self._enabled = True
        _CACHES.append(self)

    def set_enabled(self, enabled):
        self._enabled = bool(enabled)

    @property
    def enabled(self):
        return self._enabled

    def _acquire(self, key):
        """(workspace, lock) with the key's lock already held, or (None, None).
        The lock is taken inside the guard so clear() cannot release the
        workspace between lookup and use."""
        if not self._enabled:
            return None, None
        with self._guard:
            if not self._enabled:
                return None, None
            item = self._items.get(key)
            if item is None:
                try:
                    ws = self._factory()
                except Exception:
                    self._enabled = False
                    return None, None
                item = (ws, threading.Lock())
                self._items[key] = item
            ws, lock = item
            lock.acquire()
        return ws, lock

    @contextmanager
    def use(self, key):
        ws, lock = self._acquire(key)
        if lock is None:
            yield None
            return
        try:
            yield ws
        finally:
            lock.release()

    def clear(self):
        with self._guard:
            items = list(self._items.values())
            self._items = {}
            self._enabled = True
        for ws, lock in items:
            with lock:
                ws.clear()
