"""Process teardown with native objects still alive.

A failing pytest test keeps its frames (and the contexts, bases, plans, Fock
builders and servers they hold) alive in sys.last_traceback until the
interpreter exits.  The library must then release them in order and exit
cleanly: the Python atexit hook frees every live handle before the modules
are torn down, and the library's own exit handler stops server threads and
turns later destroy calls into no-ops before the HIP runtime goes."""
import os
import subprocess
import sys

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

SCRIPT = r"""
import sys
sys.path.insert(0, {root!r})
import numpy as np
import libint_amd as la
from libint_amd import molecule

mol = molecule.make_molecule(2, natoms=2)
ctx = la.Context(0)
b = la.Basis(ctx, mol.basis)
q = molecule.unique_quartets(mol.basis.nshell)[:2000]
plan = la.EriPlan(b, q)
plan.execute()
fb = la.FockBuilder(b)
fb.build(np.eye(mol.nbf) * 0.1, 1e-12)
eng = la.Engine(6, 2)               # the shared default context
s = la.Shell(1, [(1.0, 1.0)])
eng.compute(s, s, s, s)
srv = la.Server(0, hbm_bytes=2 ** 30)
rid = srv.submit(mol.basis, np.eye(mol.nbf) * 0.1, 1e-12)   # never collected
other = la.Context(0)               # a second context, never closed
ob = la.Basis(other, mol.basis)
print("objects alive", flush=True)
{tail}
"""

KEEP = r"""
def failing():
    local_plan = plan
    raise RuntimeError("planted failure")
try:
    failing()
except RuntimeError:
    # what pytest does for a failed test: the frames stay reachable
    sys.last_type, sys.last_value, sys.last_traceback = sys.exc_info()
print("END", flush=True)
"""

RAISE = r"""
print("END", flush=True)
raise RuntimeError("planted failure, uncaught")
"""


@pytest.mark.parametrize("tail,rc", [(KEEP, 0), (RAISE, 1)], ids=["traceback_kept", "uncaught"])
def test_exit_with_live_objects(tail, rc):
    print(f"case: teardown with live objects, {'uncaught exception' if rc else 'traceback kept'}", flush=True)
    p = subprocess.run([sys.executable, "-c", SCRIPT.format(root=ROOT, tail=tail)], capture_output=True, text=True,
                       timeout=180)
    print(p.stdout[-2000:], p.stderr[-2000:], flush=True)
    assert "END" in p.stdout
    assert p.returncode == rc
