"""CPU oracle -- test infrastructure only (ctypes bindings in oracle.oracle)."""
