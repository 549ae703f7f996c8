"""Removable attribute patches, kept as a stack of undo actions.

The contract is the one PySCF users of the reference rely on
(core/hooks.py:11-60 of the reference): ``wrap(owner, name, factory)``
puts ``factory(current)`` in place of ``owner.name``; ``restore_all()``
undoes every patch and every registered callback, last first.  Undoing a
patch writes back the value the owner itself held; when the owner held
none (an instance whose attribute came from its class), the patch is
deleted so the class attribute shows through again.

The stack is a ``contextlib.ExitStack``: each patch pushes the closure that
reverts it, and closing the stack runs them in reverse order, also when one
of them raises (the remaining ones still run, the first error propagates).
"""
import contextlib


def _reverter(owner, name):
    """The action that returns owner.name to its present state."""
    table = getattr(owner, '__dict__', None)
    if table is not None and name not in table:
        # nothing of its own: the patch will only shadow an inherited value
        return lambda: delattr(owner, name)
    held = table[name] if table is not None else getattr(owner, name)
    return lambda: setattr(owner, name, held)


class HookSet:
    """Patches applied to PySCF objects, undone together."""

    def __init__(self):
        self._undo = contextlib.ExitStack()
        self.closed = False

    def _check_open(self):
        if self.closed:
            raise RuntimeError('this HookSet was restored; make a new one')

    def wrap(self, owner, name, factory):
        """owner.name = factory(owner.name) until restore_all()."""
        self._check_open()
        revert = _reverter(owner, name)
        replacement = factory(getattr(owner, name))
        setattr(owner, name, replacement)
        self._undo.callback(revert)
        return replacement

    def on_restore(self, fn):
        """Run fn() at restore_all(), in stack order with the patches."""
        self._check_open()
        self._undo.callback(fn)

    def restore_all(self):
        self.closed = True
        self._undo.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.restore_all()
        return False
