"""Point the dml package at the CPU SIMT build of libdml (tests/simt/libdml_sim.so).

Test infrastructure only: the package itself loads nothing but the gfx950
libdml.so.  Processes that load the AddressSanitizer build must start with
libasan preloaded (see env())."""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
LIB = os.environ.get("SIMT_LIB") or os.path.join(HERE, "_build", "libdml_sim.so")


def build(asan=True):
    env = dict(os.environ, SIMT_ASAN="1" if asan else "0")
    subprocess.check_call(["bash", os.path.join(HERE, "build.sh")], env=env)


def env(asan=True):
    """environment for a child process that loads the sanitized build"""
    e = dict(os.environ)
    if asan:
        lib = subprocess.check_output(["g++", "-print-file-name=libasan.so"]).decode().strip()
        e["LD_PRELOAD"] = lib
        e["ASAN_OPTIONS"] = ("detect_leaks=0:malloc_fill_byte=190:max_malloc_fill_size=1073741824:"
                             "abort_on_error=1:detect_stack_use_after_return=0")
    e["PYTHONPATH"] = os.pathsep.join([os.path.join(ROOT, "din-mol-li_amd"),
                                       os.path.join(ROOT, "oracle"), e.get("PYTHONPATH", "")])
    return e


def use():
    """make `import dml` load the CPU build; returns the module"""
    for p in (os.path.join(ROOT, "din-mol-li_amd"), os.path.join(ROOT, "oracle")):
        if p not in sys.path:
            sys.path.insert(0, p)
    import dml
    dml.LIB_PATH = LIB
    dml._lib = None
    dml.lib()
    return dml
