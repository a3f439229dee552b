"""Oracle package (test infrastructure only): CPU FP64 restatement of pmsim's solid path."""
