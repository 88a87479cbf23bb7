"""Host logic of the ctypes binding (no GPU): handles made on other handles
are destroyed before them, whichever is closed or finalised first."""
from quickcbc_amd.qcbc import _Handle


class _Fake(_Handle):
    def __init__(self, name, log, *owners):
        self.name, self.log, self.h = name, log, 1
        self._depend_on(*owners)

    def close(self):
        self._close_deps()
        if self.h:
            self.log.append(self.name)
            self.h = None
        self._release()


def test_owner_closes_its_dependents_first():
    log = []
    ctx = _Fake("ctx", log)
    het = _Fake("het", log, ctx)
    pt = _Fake("pt", log, ctx, het)
    pt2 = _Fake("pt2", log, ctx)
    pt2.close()
    ctx.close()
    assert log == ["pt2", "pt", "het", "ctx"]
    assert not ctx.__dict__.get("_deps") and not het.__dict__.get("_deps")


def test_a_dependent_closed_first_is_not_closed_again():
    log = []
    ctx = _Fake("ctx", log)
    pt = _Fake("pt", log, ctx)
    pt.close()
    pt.close()
    ctx.close()
    assert log == ["pt", "ctx"]
