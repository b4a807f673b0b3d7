"""Host-side Python bindings of libmcsim.so (the MI355X milling-physics library)."""
