"""Running a spiffe input deck, and reading SDDS files, over the C ABI
(spf_deck_run, spf_sdds_*).  The deck runner itself is C++ in the library;
`python -m spiffe_amd.deck <deck>` and the `spiffe-amd` executable are two
front ends to it."""
import ctypes
import sys

import numpy as np

from . import abi


class DeckError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("spiffe_amd deck error %d: %s" % (code, msg))
        self.code = code


def run_deck(path, max_particles=0, max_steps=0, compat=True, verbose=False, device=0):
    """Runs the deck; returns the run's counts (steps, particles, emitted, screen_rows, ...)."""
    lib = abi.load()
    o = abi.DeckOptions()
    lib.spf_deck_defaults(ctypes.byref(o))
    o.device, o.compat, o.verbose = device, 1 if compat else 0, 1 if verbose else 0
    o.max_particles, o.max_steps = int(max_particles), int(max_steps)
    st = abi.DeckStats()
    err = ctypes.create_string_buffer(1024)
    rc = lib.spf_deck_run(str(path).encode(), ctypes.byref(o), ctypes.byref(st), err, len(err))
    if rc:
        raise DeckError(rc, err.value.decode())
    return {n: getattr(st, n) for n, _ in abi.DeckStats._fields_}


def sdds_info(path):
    lib = abi.load()
    n, r = ctypes.c_int64(), ctypes.c_int64()
    err = ctypes.create_string_buffer(512)
    if lib.spf_sdds_info(str(path).encode(), ctypes.byref(n), ctypes.byref(r), err, len(err)):
        raise ValueError(err.value.decode())
    return n.value, r.value


def sdds_column(path, column, page=0):
    lib = abi.load()
    n = ctypes.c_int64()
    err = ctypes.create_string_buffer(512)
    if lib.spf_sdds_column(str(path).encode(), page, column.encode(), None, 0, ctypes.byref(n), err, len(err)):
        raise ValueError(err.value.decode())
    out = np.empty(n.value)
    if lib.spf_sdds_column(str(path).encode(), page, column.encode(), out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                           n.value, ctypes.byref(n), err, len(err)):
        raise ValueError(err.value.decode())
    return out


def sdds_parameter(path, name, page=0):
    lib = abi.load()
    v = ctypes.c_double()
    err = ctypes.create_string_buffer(512)
    if lib.spf_sdds_parameter(str(path).encode(), page, name.encode(), ctypes.byref(v), err, len(err)):
        raise ValueError(err.value.decode())
    return v.value


if __name__ == "__main__":
    if len(sys.argv) < 2:
        print("usage: python -m spiffe_amd.deck <deck> [max_particles] [max_steps]")
        sys.exit(1)
    stats = run_deck(sys.argv[1], int(sys.argv[2]) if len(sys.argv) > 2 else 0,
                     int(sys.argv[3]) if len(sys.argv) > 3 else 0, verbose=True)
    print(stats)
