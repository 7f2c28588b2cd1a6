"""The oracle: CPU FP64 restatement of the SPM step. TEST INFRASTRUCTURE ONLY."""
