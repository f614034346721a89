"""MI355X engine for the sim-trhepd-rheed rocking-curve slice loop (host-side package).

deckio  -- bulk.txt / surf.txt / bulkP.b / .s formats
lib     -- ctypes binding of the C ABI (libtrhepd_hip.so)
"""
from . import deckio  # noqa: F401
