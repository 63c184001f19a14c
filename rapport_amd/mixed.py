"""Mixed-precision path (complex64 field, split-FP16 Hankel GEMM, FP64 plasma scan)."""
raise NotImplementedError("precision='mixed' is not built yet")
