"""CPU oracle for the maximal-biclique enumeration — TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg use this
package, as the checker; the engine (gmbe_amd) never imports it.
"""
from .binding import (OracleResult, available, brute, gmbe, imbea, mbea, pmbe,  # noqa: F401
                      hash_pair, LIB_PATH)
