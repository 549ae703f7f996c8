"""Reversible attribute patching (the reference's core/hooks.py contract).

``HookSet.wrap(owner, name, factory)`` replaces ``owner.name`` by
``factory(original)``; ``restore_all()`` undoes every patch in reverse
order.  An attribute that was defined on the owner itself (a module
function, a class attribute) is written back; one that only shadowed a
class-level attribute on an instance is deleted, re-exposing the class one
(core/hooks.py:11-23 of the reference).
"""
_MISSING = object()


def _install(owner, name, replacement):
    own = vars(owner).get(name, _MISSING) if hasattr(owner, '__dict__') else _MISSING
    setattr(owner, name, replacement)
    return (owner, name, own)


def _restore(token):
    owner, name, own = token
    if own is _MISSING:
        delattr(owner, name)
    else:
        setattr(owner, name, own)


class HookSet:
    def __init__(self):
        self._tokens = []
        self._done = False

    def wrap(self, owner, name, factory):
        if self._done:
            raise RuntimeError('HookSet already restored')
        token = _install(owner, name, factory(getattr(owner, name)))
        self._tokens.append(token)
        return token

    def on_restore(self, fn):
        if self._done:
            raise RuntimeError('HookSet already restored')
        self._tokens.append(fn)

    def restore_all(self):
        for token in reversed(self._tokens):
            if callable(token):
                token()
            else:
                _restore(token)
        self._tokens.clear()
        self._done = True

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.restore_all()
