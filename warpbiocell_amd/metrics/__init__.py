"""Population metrics with the reductions on the device (mirror of the reference's metrics package)."""
