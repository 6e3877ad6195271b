"""Context teardown order (no GPU): a context left open is destroyed by the
atexit hook, before interpreter finalization; a finalizer that runs during
finalization never calls into the library; close() is idempotent.  The
library is replaced by a recorder, so nothing here touches a device."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

SCRIPT = r"""
import sys
sys.path.insert(0, {root!r})
from idp_amd import capi

class Fake:
    def __init__(self):
        self.next = 1
    def idp_ctx_create(self, dev, out):
        out._obj.value = self.next
        self.next += 1
        return 0
    def idp_ctx_destroy(self, h):
        print("destroy", h.value, "finalizing" if sys.is_finalizing() else "live", flush=True)
        return 0
    def idp_set_launch_config(self, *a):
        return 0
    def idp_last_error(self, h):
        return b""

capi._lib = Fake()
a = capi.Context(0)
b = capi.Context(0)
c = capi.Context(0)
c.close()
c.close()
cycle = [b]
cycle.append(cycle)  # b lives in a reference cycle until the end
del b, cycle
print("end of script", flush=True)
"""


def test_open_contexts_are_destroyed_once_before_finalization():
    p = subprocess.run([sys.executable, "-c", SCRIPT.format(root=ROOT)], capture_output=True, text=True,
                       timeout=120)
    assert p.returncode == 0, p.stderr
    lines = p.stdout.strip().splitlines()
    destroyed = [ln for ln in lines if ln.startswith("destroy")]
    assert sorted(destroyed) == ["destroy 1 live", "destroy 2 live", "destroy 3 live"], lines
    # the explicitly closed one went first, the others at exit
    assert lines[0] == "destroy 3 live" and lines[1] == "end of script"
