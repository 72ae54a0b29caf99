"""FP64 CPU oracle for the FastFFT path (test infrastructure, not product)."""
