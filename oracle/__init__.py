"""TEST INFRASTRUCTURE -- the oracle: a CPU restatement of the reference step loop."""
