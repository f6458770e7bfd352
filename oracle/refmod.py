"""TEST INFRASTRUCTURE: loader of the reference native gridder built from its own sources.

`make -f oracle/Makefile.ref` compiles DDFacet/Gridder/{GridderSmearPols,Semaphores,JonesServer,
DecorrelationHelper}.cc into oracle/_ref/ (git-ignored).  This module imports that build, nothing
else of the reference.  The degridder takes a named POSIX semaphore per row group
(Semaphores.cc:27-49), so ``module()`` registers a small private semaphore set first.
"""
import atexit
import glob
import importlib.util
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
_MOD = None


def available():
    return bool(glob.glob(os.path.join(_HERE, "_ref", "_pyGridderSmearPols3x*.so")))


def module(nsem=7):
    """The reference module with `nsem` row semaphores registered (the reference deployment
    uses NSemaphores = 3373, ClassFacetMachine.py:156-176); the first call fixes the set."""
    global _MOD
    if _MOD is None:
        paths = glob.glob(os.path.join(_HERE, "_ref", "_pyGridderSmearPols3x*.so"))
        if not paths:
            raise RuntimeError("reference module not built: make -f oracle/Makefile.ref")
        spec = importlib.util.spec_from_file_location("_pyGridderSmearPols3x", paths[0])
        mod = importlib.util.module_from_spec(spec)
        spec.loader.exec_module(mod)
        names = ["/ddfgpu_ref_%d_%d" % (os.getpid(), i) for i in range(int(nsem))]
        mod.pySetSemaphores(names)
        atexit.register(mod.pyDeleteSemaphore)
        _MOD = mod
    return _MOD
